"""hmatrixgpu -- Python mirror of the HMatrixGPU.jl API over libhmx (MI355X).

The H-matrix product runs in hand-written gfx950 HIP kernels; cluster tree,
block tree and ACA+/SVD compression run natively in C++.  See trees.py and
hmatrix.py for the reference file:line each name mirrors.
"""
from ._lib import HMXError, HMX_BF16, HMX_F32, HMX_F64, LIB_PATH  # noqa: F401
from .trees import (ACA_plus, BlockTree, BlockTreeNode, ClusterNode, ClusterTree,  # noqa: F401
                    SVD_recompress, info, merge_dense_matrices_, traverse)
from .hmatrix import (DipoleKernel, HMatrix, HMatrixBatch, HMatrixCPU, LaplaceKernel,  # noqa: F401
                      Log2DKernel,
                      RankPredictor, mul_, product, set_backend, set_groupsize, sparsify_hmatrix)

__all__ = ["ClusterTree", "ClusterNode", "BlockTree", "BlockTreeNode", "merge_dense_matrices_",
           "traverse", "ACA_plus", "SVD_recompress", "HMatrix", "HMatrixCPU", "LaplaceKernel",
           "Log2DKernel", "mul_", "product", "info", "set_backend", "set_groupsize", "HMXError",
           "HMatrixBatch", "RankPredictor", "sparsify_hmatrix", "DipoleKernel"]
