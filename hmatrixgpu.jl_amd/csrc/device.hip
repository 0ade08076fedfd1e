// device.hip -- gfx950 device layout, panel generation and the H-matrix product.
//
// Reference counterpart: the two KernelAbstractions kernels and the driver of
// mult.jl:41-189 (V_mult_vec_kernel!, DU_mult_vec_kernel!, mul!).  The work is
// reorganised for CDNA4 instead of translated:
//
//   * Leaf panels.  Every output row belongs to exactly one target-tree leaf.
//     For each leaf (split into row groups of <= 64 rows) all its near-field
//     dense-block rows and all its low-rank U rows are packed into ONE panel
//     of `rows x K` entries, K = sum of the dense blocks' column counts plus
//     the ranks of the low-rank blocks covering the leaf.  Column c of the
//     panel multiplies row colidx[c] of a combined source array
//         S = [ X permuted by source_index_map (n rows) ; T = V*X (R rows) ; 0 ].
//     The product of a leaf is then one small GEMM  Y_leaf = P_leaf * S[colidx]
//     written once, with no atomics and no O(m * blocks) scan (the reference's
//     DU kernel tests every block for every row, mult.jl:95-118).
//   * Far field.  V rows are grouped by source cluster (all blocks sharing a
//     source cluster read the same contiguous S rows) into panels of <= 64
//     rank-rows; long clusters are split into column chunks whose partial
//     results are summed in a fixed order (deterministic, atomic-free).
//   * Tiles.  A panel is stored as k-tiles of TK = 64 bytes / sizeof(storage)
//     columns; inside a k-tile, row tiles of 16 rows are stored chunk-major so
//     that lane l of a wave reads 16 contiguous bytes = row (l % 16), chunk
//     (l / 16): one fully coalesced 1 KiB load per wave per 16x TK tile, which
//     is exactly the A-operand layout of MFMA 16x16x4 (f32/f64) and 16x16x32
//     (bf16).  Partial row tiles are stored compactly (no padding bytes).
//   * Multi-RHS (16 columns) runs on MFMA with the RHS as the N dimension;
//     a single RHS is a VALU GEMV over the same tiles.
//
// Stages per call (one launch each, stream ordered): gather X -> S, far-field
// panels (V*X), partial reduce, leaf panels (near field + U*T) -> scatter to Y.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "hmx_host.hpp"

namespace hmx {

#define HIPCHECK(expr)                                                                       \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            throw Error(HMX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
    } while (0)

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct bf16s {
    uint16_t u;
};

// One panel work item (rows <= 64).
struct Item {
    int64_t data_off;  // element offset of the panel in the storage array
    int64_t col_off;   // leaf panels: offset into colidx; far-field: first S row
    int64_t out_off;   // leaf panels: first permuted target row; far-field: output row
    int32_t rows;
    int32_t ktiles;
    int32_t ncols;     // far-field: valid columns (the rest of the last tile is masked)
    int32_t to_T;      // far-field: 1 = write T rows of S directly, 0 = partial rows
};

struct ReduceGroup {
    int64_t pbase;  // first partial row
    int64_t tbase;  // first T row (S row = n + tbase)
    int32_t rows;
    int32_t nch;
};

// ---- element conversions ---------------------------------------------------
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return static_cast<uint16_t>(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

template <typename T>
__device__ inline T from_double(double v);
template <>
__device__ inline double from_double<double>(double v) { return v; }
template <>
__device__ inline float from_double<float>(double v) { return static_cast<float>(v); }
template <>
__device__ inline bf16s from_double<bf16s>(double v) { return bf16s{f2bf(static_cast<float>(v))}; }

template <typename T>
__device__ inline T from_float(float v);
template <>
__device__ inline double from_float<double>(float v) { return v; }
template <>
__device__ inline float from_float<float>(float v) { return v; }
template <>
__device__ inline bf16s from_float<bf16s>(float v) { return bf16s{f2bf(v)}; }

// ---- storage traits ----------------------------------------------------------
// TK: columns per k-tile (64 bytes per row), EPL: elements per lane chunk.
template <typename TS>
struct Store;
template <>
struct Store<double> {
    static constexpr int TK = 8, EPL = 2;
    using acc = double;
};
template <>
struct Store<float> {
    static constexpr int TK = 16, EPL = 4;
    using acc = float;
};
template <>
struct Store<bf16s> {
    static constexpr int TK = 32, EPL = 8;
    using acc = float;
};
// B-operand (S) element type: W = 16 columns -> MFMA operand type, W = 1 -> acc
template <typename TS, int W>
struct SType {
    using type = typename Store<TS>::acc;
};
template <>
struct SType<double, 16> {
    using type = double;
};
template <>
struct SType<float, 16> {
    using type = float;
};
template <>
struct SType<bf16s, 16> {
    using type = bf16s;
};

constexpr int kWaves = 4;  // waves per workgroup in the panel kernels

// ---- the panel kernel ----------------------------------------------------------
// Buffer resources (raw, stride 0): loads past num_records return 0, which masks
// the absent rows of a partial row tile, the padding of the last k-tile and the
// prefetch past the end of a panel without branches.
constexpr int kRsrcWord3 = 0x00020000;  // gfx9-family raw buffer descriptor
// Cache policy of the streamed panel loads: non-temporal, so that the matrix
// stream (read once per product) does not evict the reused source / T rows
// from L2 (measured: 3-10% faster than the default policy at N = 262,144).
constexpr int kPanelPolicy = 2;
constexpr uint32_t kOOB = 0xffffffffu;

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes),
                                             kRsrcWord3);
}

template <int N>
struct IdxV {
    int32_t v[N];
};

// Column indices of the lane's EPL columns of k-tile kt.
template <int EPL, int TK>
__device__ inline IdxV<EPL> load_idx(__amdgpu_buffer_rsrc_t C, int kt, int lg) {
    IdxV<EPL> r;
    const uint32_t off = static_cast<uint32_t>((kt * TK + lg * EPL) * 4);
    if constexpr (EPL == 2) {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(C, off, 0, 0);
        r.v[0] = static_cast<int32_t>(v[0]);
        r.v[1] = static_cast<int32_t>(v[1]);
    } else if constexpr (EPL == 4) {
        auto v = __builtin_amdgcn_raw_buffer_load_b128(C, off, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; e++) r.v[e] = static_cast<int32_t>(v[e]);
    } else {
        auto v0 = __builtin_amdgcn_raw_buffer_load_b128(C, off, 0, 0);
        auto v1 = __builtin_amdgcn_raw_buffer_load_b128(C, off + 16, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            r.v[e] = static_cast<int32_t>(v0[e]);
            r.v[4 + e] = static_cast<int32_t>(v1[e]);
        }
    }
    return r;
}

template <typename T>
__device__ inline T buf_load(__amdgpu_buffer_rsrc_t R, uint32_t off);
template <>
__device__ inline double buf_load<double>(__amdgpu_buffer_rsrc_t R, uint32_t off) {
    auto v = __builtin_amdgcn_raw_buffer_load_b64(R, off, 0, 0);
    return __builtin_bit_cast(double, v);
}
template <>
__device__ inline float buf_load<float>(__amdgpu_buffer_rsrc_t R, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(R, off, 0, 0));
}
template <>
__device__ inline bf16s buf_load<bf16s>(__amdgpu_buffer_rsrc_t R, uint32_t off) {
    return bf16s{__builtin_amdgcn_raw_buffer_load_b16(R, off, 0, 0)};
}

// One wave's share of a panel: k-tiles kt0, kt0 + kstep, ... of a `rows` x
// (ktiles * TK) panel at pdata.  B rows come from S: through the column index
// array (LEAF) or as the contiguous rows col_off + k masked beyond ncols.  The
// accumulated results are handed to emit(row, col, value), row < 64.  Loads are
// issued one k-tile ahead (A and B) and two ahead (column indices).  CP is the
// cache policy of the streamed A loads.
template <typename TS, int W, bool LEAF, int NRT, int CP, typename Emit>
__device__ __forceinline__ void panel_body(const TS* pdata, int rows, int ktiles, int64_t col_off,
                                           int ncols, const int32_t* colidx,
                                           const typename SType<TS, W>::type* S, uint32_t s_bytes,
                                           int kt0, int kstep, int lane, Emit&& emit) {
    using acc_t = typename Store<TS>::acc;
    using s_t = typename SType<TS, W>::type;
    constexpr int TK = Store<TS>::TK, EPL = Store<TS>::EPL;
    constexpr uint32_t SB = sizeof(s_t), ROWB = W * sizeof(s_t);
    const int li = lane & 15, lg = lane >> 4;
    const uint32_t slab = static_cast<uint32_t>(rows) * 64u;
    const __amdgpu_buffer_rsrc_t Ar = make_rsrc(pdata, slab * ktiles);
    const __amdgpu_buffer_rsrc_t Sr = make_rsrc(S, s_bytes);
    const __amdgpu_buffer_rsrc_t Cr =
        make_rsrc(LEAF ? colidx + col_off : colidx, LEAF ? ktiles * TK * 4 : 0);
    uint32_t aoff[NRT];
    bool aval[NRT];
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) {
        const int rr = min(16, rows - 16 * rt);
        aval[rt] = li < rr;
        aoff[rt] = static_cast<uint32_t>((16 * rt * TK + (rr * lg + li) * EPL) * sizeof(TS));
    }
    const uint32_t bcol = (W == 16 ? li : 0) * SB;

    struct Stage {
        u32x4 a[NRT];
        s_t b[EPL];
    };
    auto load_stage = [&](int kt, const IdxV<EPL>& ix) {
        Stage st;
        const uint32_t base = static_cast<uint32_t>(kt) * slab;
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) {
            const uint32_t o = aval[rt] ? base + aoff[rt] : kOOB;
            st.a[rt] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(Ar, o, 0, CP));
        }
#pragma unroll
        for (int e = 0; e < EPL; e++) {
            uint32_t o;
            if constexpr (LEAF) {
                o = static_cast<uint32_t>(ix.v[e]) * ROWB + bcol;
            } else {
                const int k = kt * TK + lg * EPL + e;
                o = k < ncols ? static_cast<uint32_t>(col_off + k) * ROWB + bcol : kOOB;
            }
            st.b[e] = buf_load<s_t>(Sr, o);
        }
        return st;
    };

    using vacc = typename std::conditional<W == 16,
        typename std::conditional<std::is_same<acc_t, double>::value, f64x4, f32x4>::type,
        acc_t>::type;
    vacc acc[NRT];
#pragma unroll
    for (int rt = 0; rt < NRT; rt++) acc[rt] = vacc{};

    auto compute = [&](const Stage& st) {
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) {
            const u32x4 raw = st.a[rt];
            if constexpr (W == 16) {
                if constexpr (std::is_same<TS, float>::value) {
                    f32x4 a = __builtin_bit_cast(f32x4, raw);
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, st.b[0], acc[rt], 0, 0, 0);
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, st.b[1], acc[rt], 0, 0, 0);
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, st.b[2], acc[rt], 0, 0, 0);
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, st.b[3], acc[rt], 0, 0, 0);
                } else if constexpr (std::is_same<TS, double>::value) {
                    typedef double d2 __attribute__((ext_vector_type(2)));
                    d2 a = __builtin_bit_cast(d2, raw);
                    acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, st.b[0], acc[rt], 0, 0, 0);
                    acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, st.b[1], acc[rt], 0, 0, 0);
                } else {
                    bf16x8 a = __builtin_bit_cast(bf16x8, raw);
                    uint16_t bb[8];
#pragma unroll
                    for (int e = 0; e < 8; e++) bb[e] = st.b[e].u;
                    bf16x8 bv = __builtin_bit_cast(bf16x8, bb);
                    acc[rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[rt], 0, 0, 0);
                }
            } else {
                if constexpr (std::is_same<TS, float>::value) {
                    f32x4 a = __builtin_bit_cast(f32x4, raw);
                    acc[rt] += a.x * st.b[0] + a.y * st.b[1] + a.z * st.b[2] + a.w * st.b[3];
                } else if constexpr (std::is_same<TS, double>::value) {
                    typedef double d2 __attribute__((ext_vector_type(2)));
                    d2 a = __builtin_bit_cast(d2, raw);
                    acc[rt] += a.x * st.b[0] + a.y * st.b[1];
                } else {
                    uint16_t h[8];
                    std::memcpy(h, &raw, 16);
                    acc_t s = 0;
#pragma unroll
                    for (int e = 0; e < 8; e++) s += bf2f(h[e]) * st.b[e];
                    acc[rt] += s;
                }
            }
        }
    };

    // Software pipeline, unrolled by two so the stage buffers alternate without
    // register copies: while stage A is consumed, stage B (next k-tile) and the
    // column indices of the k-tile after it are in flight.  Loads past the end
    // of the panel (kt >= ktiles) fall outside the buffer and return zeros;
    // they are never consumed.
    int kt = kt0;
    IdxV<EPL> ixA{}, ixB{};
    if constexpr (LEAF) {
        ixA = load_idx<EPL, TK>(Cr, kt, lg);
        ixB = load_idx<EPL, TK>(Cr, kt + kstep, lg);
    }
    Stage sA = load_stage(kt, ixA), sB;
    while (kt < ktiles) {
        if constexpr (LEAF) ixA = load_idx<EPL, TK>(Cr, kt + 2 * kstep, lg);
        sB = load_stage(kt + kstep, ixB);
        compute(sA);
        kt += kstep;
        if (kt >= ktiles) break;
        if constexpr (LEAF) ixB = load_idx<EPL, TK>(Cr, kt + 2 * kstep, lg);
        sA = load_stage(kt + kstep, ixA);
        compute(sB);
        kt += kstep;
    }

    if constexpr (W == 16) {
        // MFMA 16x16 output, column li: f32 accumulators hold rows 4*lg+v of the
        // row tile, the f64 instruction holds rows 4*v+lg.
#pragma unroll
        for (int rt = 0; rt < NRT; rt++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = std::is_same<TS, double>::value ? 4 * v + lg : 4 * lg + v;
                emit(16 * rt + row, li, static_cast<acc_t>(acc[rt][v]));
            }
    } else {
#pragma unroll
        for (int rt = 0; rt < NRT; rt++) {
            acc_t v = acc[rt];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (lg == 0) emit(16 * rt + li, 0, v);
        }
    }
}

template <typename F>
__device__ __forceinline__ void dispatch_nrt(int rows, F&& f) {
    switch ((rows + 15) >> 4) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    default: f(std::integral_constant<int, 4>{}); break;
    }
}

// Leaf panel: the kWaves waves split the k-tiles of one item; partial sums are
// reduced through LDS and scattered to Y through target_map.
template <typename TS, int W, int CP, typename TIO>
__global__ __launch_bounds__(kWaves * 64) void k_leaf(
    const Item* __restrict__ items, const TS* __restrict__ panel,
    const int32_t* __restrict__ colidx, const typename SType<TS, W>::type* __restrict__ S,
    uint32_t s_bytes, const int32_t* __restrict__ tmap0, TIO* __restrict__ Y, int64_t ldy,
    int32_t rhs0, int32_t nrhs) {
    using acc_t = typename Store<TS>::acc;
    const Item it = items[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rows = it.rows;
    __shared__ acc_t red[kWaves][64][W];
    auto emit = [&](int row, int j, acc_t v) { red[wave][row][j] = v; };
    dispatch_nrt(rows, [&](auto nrt) {
        panel_body<TS, W, true, decltype(nrt)::value, CP>(panel + it.data_off, rows, it.ktiles,
                                                          it.col_off, it.ncols, colidx, S, s_bytes,
                                                          wave, kWaves, lane, emit);
    });
    __syncthreads();
    for (int t = threadIdx.x; t < rows * W; t += kWaves * 64) {
        const int r = t / W, j = t % W;
        acc_t s = red[0][r][j];
#pragma unroll
        for (int w = 1; w < kWaves; w++) s += red[w][r][j];
        if (rhs0 + j < nrhs)
            Y[static_cast<int64_t>(tmap0[it.out_off + r]) + static_cast<int64_t>(rhs0 + j) * ldy] =
                static_cast<TIO>(s);
    }
}

// Far-field work item: up to kWaves consecutive 64-row groups of the stacked V
// rows of one source cluster over one column chunk.  Wave w owns row group w
// and walks every k-tile of the chunk; the waves read the same S rows (the
// cluster's X rows), so the B operand is fetched once per workgroup and shared
// through the L1 cache.  No cross-wave reduction is needed.
struct FarItem {
    int64_t data_off;  // panel of row group 0; group g at + g*64*ktiles*TK
    int64_t col_off;   // first S row of the chunk
    int64_t out_off;   // output row of row 0 of group 0 (group g: + 64*g)
    int32_t ktiles;
    int32_t ncols;
    int32_t rows_last;  // rows of the last group (the others have 64)
    int16_t nrg;        // number of row groups (<= kWaves)
    int16_t to_T;       // 1: write T rows of S, 0: write partial rows
};

template <typename TS, int W, int CP>
__global__ __launch_bounds__(kWaves * 64) void k_far(
    const FarItem* __restrict__ items, const TS* __restrict__ panel,
    typename SType<TS, W>::type* __restrict__ S, int64_t nsrc, uint32_t s_bytes,
    typename Store<TS>::acc* __restrict__ partial) {
    using acc_t = typename Store<TS>::acc;
    using s_t = typename SType<TS, W>::type;
    constexpr int TK = Store<TS>::TK;
    const FarItem it = items[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave >= it.nrg) return;
    const int rows = wave == it.nrg - 1 ? it.rows_last : 64;
    const TS* pd = panel + it.data_off + static_cast<int64_t>(wave) * 64 * it.ktiles * TK;
    const int64_t orow = it.out_off + 64 * wave;
    auto emit = [&](int row, int j, acc_t v) {
        if (row >= rows) return;
        if (it.to_T)
            S[(nsrc + orow + row) * W + j] = from_double<s_t>(static_cast<double>(v));
        else
            partial[(orow + row) * W + j] = v;
    };
    dispatch_nrt(rows, [&](auto nrt) {
        panel_body<TS, W, false, decltype(nrt)::value, CP>(pd, rows, it.ktiles, it.col_off,
                                                           it.ncols, nullptr, S, s_bytes, 0, 1,
                                                           lane, emit);
    });
}

// Sum the column-chunk partials of far-field row groups in a fixed order.
template <typename TS, int W>
__global__ __launch_bounds__(256) void k_reduce(const ReduceGroup* __restrict__ groups,
                                                const typename Store<TS>::acc* __restrict__ partial,
                                                typename SType<TS, W>::type* __restrict__ S,
                                                int64_t nsrc) {
    using acc_t = typename Store<TS>::acc;
    using s_t = typename SType<TS, W>::type;
    const ReduceGroup g = groups[blockIdx.x];
    for (int t = threadIdx.x; t < g.rows * W; t += blockDim.x) {
        const int r = t / W, j = t % W;
        acc_t s = 0;
        for (int c = 0; c < g.nch; c++) s += partial[(g.pbase + static_cast<int64_t>(c) * g.rows + r) * W + j];
        S[(nsrc + g.tbase + r) * W + j] = from_double<s_t>(static_cast<double>(s));
    }
}

// S[p][j] = X[smap[p], rhs0 + j]  (zero for j beyond nrhs); thread per original index.
template <typename s_t, int W, typename TIO>
__global__ __launch_bounds__(256) void k_gather(const TIO* __restrict__ X, int64_t ldx,
                                                const int32_t* __restrict__ inv_smap, int64_t n,
                                                s_t* __restrict__ S, int32_t rhs0, int32_t nrhs) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t p = inv_smap[t];
#pragma unroll
    for (int j = 0; j < W; j++) {
        const int c = rhs0 + j;
        const double v = c < nrhs ? static_cast<double>(X[t + static_cast<int64_t>(c) * ldx]) : 0.0;
        S[p * W + j] = from_double<s_t>(v);
    }
}

// ---- build-time panel generation ------------------------------------------------
// Column descriptors of the leaf panels: kind 0 = built-in kernel entry
// (base = 0-based source point), 1 = U payload, 2 = dense payload (callback
// kernels), 3 = padding.  Buffer element (row r) = base + r * stride.
struct ColDesc {
    int64_t base;
    int32_t stride;
    int32_t kind;
};

__device__ inline double kernel_entry(int kid, const double* a, const double* b) {
#pragma clang fp contract(off)
    if (kid == HMX_KERNEL_LAPLACE) {
        const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
        double d2 = dx * dx + dy * dy;
        d2 = d2 + dz * dz;
        const double d = sqrt(d2);
        return d == 0.0 ? 0.0 : 1.0 / d;
    }
    // HMX_KERNEL_LOG2D: generic_norm2 branch without scaling (scaling only
    // triggers for |r| < 1e-154, where the entry is 1.0 anyway).
    const double r0 = a[0] - b[0], r1 = a[1] - b[1], r2 = a[2] - b[2];
    double s = r0 * r0;
    s += r1 * r1;
    s += r2 * r2;
    const double d = sqrt(s);
    return d < 1e-12 ? 1.0 : (-0.5 / 3.141592653589793) * log(d);
}

// Element (row r, column k) of an item, at tile position q (16-byte chunks).
template <typename TS>
__device__ inline void chunk_coords(int q, int rows, int& kt, int& r, int& k0) {
    constexpr int TK = Store<TS>::TK, EPL = Store<TS>::EPL;
    const int per_slab = rows * 4;
    kt = q / per_slab;
    const int c = q % per_slab;
    const int rt = c / 64;
    const int rr = min(16, rows - 16 * rt);
    const int w = c % 64;
    const int g = w / rr, i = w % rr;
    r = 16 * rt + i;
    k0 = kt * TK + g * EPL;
}

template <typename TS>
__global__ __launch_bounds__(256) void k_gen_leaf(const Item* __restrict__ items, TS* __restrict__ panel,
                                                  const ColDesc* __restrict__ cols,
                                                  const int32_t* __restrict__ tmap0,
                                                  const double* __restrict__ xt,
                                                  const double* __restrict__ xs,
                                                  const double* __restrict__ Ubuf,
                                                  const double* __restrict__ Dbuf, int kid) {
    constexpr int EPL = Store<TS>::EPL;
    const Item it = items[blockIdx.x];
    const int nchunks = it.ktiles * it.rows * 4;
    for (int q = threadIdx.x; q < nchunks; q += blockDim.x) {
        int kt, r, k0;
        chunk_coords<TS>(q, it.rows, kt, r, k0);
        const int64_t trow = it.out_off + r;
        const double* a = xt ? xt + 3 * static_cast<int64_t>(tmap0[trow]) : nullptr;
        TS vals[EPL];
#pragma unroll
        for (int e = 0; e < EPL; e++) {
            const ColDesc cd = cols[it.col_off + k0 + e];
            double v = 0.0;
            if (cd.kind == 0)
                v = kernel_entry(kid, a, xs + 3 * cd.base);
            else if (cd.kind == 1)
                v = Ubuf[cd.base + static_cast<int64_t>(r) * cd.stride];
            else if (cd.kind == 2)
                v = Dbuf[cd.base + static_cast<int64_t>(r) * cd.stride];
            vals[e] = from_double<TS>(v);
        }
        // chunk q of the item sits at data_off + q*EPL (chunk-major tiles)
        std::memcpy(panel + it.data_off + static_cast<int64_t>(q) * EPL, vals, sizeof(vals));
    }
}

// Far-field panels: element (r, k) = V[vrow[vbase + r] + kc0 + k] for k < ncols.
struct GenA {
    int64_t vbase;  // first entry of vrow for this item's rows
    int64_t kc0;    // first column of the chunk inside the source cluster
};

template <typename TS>
__global__ __launch_bounds__(256) void k_gen_far(const Item* __restrict__ items,
                                                 const GenA* __restrict__ gens, TS* __restrict__ panel,
                                                 const int64_t* __restrict__ vrow,
                                                 const double* __restrict__ Vbuf) {
    constexpr int EPL = Store<TS>::EPL;
    const Item it = items[blockIdx.x];
    const GenA g = gens[blockIdx.x];
    const int nchunks = it.ktiles * it.rows * 4;
    for (int q = threadIdx.x; q < nchunks; q += blockDim.x) {
        int kt, r, k0;
        chunk_coords<TS>(q, it.rows, kt, r, k0);
        const int64_t vb = vrow[g.vbase + r] + g.kc0;
        TS vals[EPL];
#pragma unroll
        for (int e = 0; e < EPL; e++) {
            const int k = k0 + e;
            vals[e] = from_double<TS>(k < it.ncols ? Vbuf[vb + k] : 0.0);
        }
        std::memcpy(panel + it.data_off + static_cast<int64_t>(q) * EPL, vals, sizeof(vals));
    }
}

// ---- host-side operator ------------------------------------------------------------
struct Scratch {
    void* S16 = nullptr;   // (n + R + 1 + pad) x 16
    void* S1 = nullptr;    // (n + R + 1 + pad)
    void* P16 = nullptr;   // partial rows x 16
    void* P1 = nullptr;
};

struct DeviceOp {
    int storage = 0, device = 0;
    int64_t m = 0, n = 0, R = 0;
    void* panel = nullptr;
    int64_t panel_elems = 0;
    Item* itemsB = nullptr;
    int64_t nB = 0;
    int32_t* colidx = nullptr;
    int64_t ncol = 0;
    FarItem* itemsA = nullptr;
    int64_t nA = 0;
    ReduceGroup* groups = nullptr;
    int64_t nG = 0;
    int64_t npartial = 0;
    int32_t* tmap0 = nullptr;
    int32_t* inv_smap = nullptr;
    std::mutex mu;
    std::map<void*, Scratch> scratch;
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[4];
    double ms[4] = {0, 0, 0, 0};
    int64_t launches[4] = {0, 0, 0, 0};
};

namespace {

size_t elem_size(int storage) { return storage == HMX_F64 ? 8 : storage == HMX_F32 ? 4 : 2; }
int tk_of(int storage) { return static_cast<int>(64 / elem_size(storage)); }

template <typename T>
T* upload(const std::vector<T>& v) {
    T* d = nullptr;
    if (v.empty()) return nullptr;
    HIPCHECK(hipMalloc(&d, v.size() * sizeof(T)));
    HIPCHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        HIPCHECK(hipGetDevice(&prev));
        if (prev != dev) HIPCHECK(hipSetDevice(dev));
    }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

// Dispatch order of the leaf items.  Default: largest first (longest
// processing time first), the best measured order at N = 262,144.
// HMX_ITEM_ORDER=xcd: the tree-ordered (spatially coherent) item list is cut
// into kXcds contiguous chunks of equal work and interleaved, so that with
// round-robin workgroup placement (workgroup i -> XCD i % kXcds) every XCD's L2
// serves one compact region; HMX_ITEM_ORDER=tree keeps tree order.
constexpr int kXcds = 8;

void order_leaf_items(std::vector<Item>& items) {
    if (items.size() < 2 * kXcds) return;
    auto work = [](const Item& it) { return int64_t(it.ktiles) * it.rows * 64 + 4096; };
    const char* env = std::getenv("HMX_ITEM_ORDER");
    if (!env || std::string(env) == "lpt") {
        std::stable_sort(items.begin(), items.end(),
                         [&](const Item& a, const Item& b) { return work(a) > work(b); });
        return;
    }
    if (env && std::string(env) == "tree") return;
    int64_t total = 0;
    for (const Item& it : items) total += work(it);
    std::vector<std::vector<Item>> chunk(kXcds);
    int64_t acc = 0;
    for (const Item& it : items) {
        const int c = static_cast<int>(std::min<int64_t>(kXcds - 1, acc * kXcds / total));
        chunk[c].push_back(it);
        acc += work(it);
    }
    std::vector<size_t> pos(kXcds, 0);
    std::vector<Item> out;
    out.reserve(items.size());
    for (size_t p = 0; out.size() < items.size(); p++) {
        int c = static_cast<int>(p % kXcds);
        if (pos[c] == chunk[c].size()) {  // exhausted: borrow from the fullest chunk
            size_t best = 0;
            for (int q = 0; q < kXcds; q++)
                if (chunk[q].size() - pos[q] > chunk[best].size() - pos[best]) best = q;
            c = static_cast<int>(best);
        }
        out.push_back(chunk[c][pos[c]++]);
    }
    items.swap(out);
}

// Far-field items: largest first (their B rows are shared inside the
// workgroup, so dispatch order only matters for the tail).
void order_far_items(std::vector<FarItem>& items) {
    std::stable_sort(items.begin(), items.end(), [](const FarItem& a, const FarItem& b) {
        return int64_t(a.ktiles) * a.nrg > int64_t(b.ktiles) * b.nrg;
    });
}

}  // namespace

void device_destroy(DeviceOp* op) {
    if (!op) return;
    (void)hipSetDevice(op->device);
    for (void* p : {op->panel, (void*)op->itemsB, (void*)op->colidx, (void*)op->itemsA,
                    (void*)op->groups, (void*)op->tmap0, (void*)op->inv_smap})
        if (p) (void)hipFree(p);
    for (auto& kv : op->scratch)
        for (void* p : {kv.second.S16, kv.second.S1, kv.second.P16, kv.second.P1})
            if (p) (void)hipFree(p);
    for (auto& e : op->ev)
        for (auto& pr : e) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    delete op;
}

int64_t device_panel_bytes(const DeviceOp* op) {
    return op ? op->panel_elems * static_cast<int64_t>(elem_size(op->storage)) : 0;
}
int64_t device_index_bytes(const DeviceOp* op) {
    if (!op) return 0;
    return op->ncol * 4 + op->nA * static_cast<int64_t>(sizeof(FarItem)) +
           op->nB * static_cast<int64_t>(sizeof(Item)) +
           op->nG * static_cast<int64_t>(sizeof(ReduceGroup)) + (op->m + op->n) * 4;
}

// Build the device operator for one storage precision.
void device_build(HMatrix& h, int storage) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        throw Error(HMX_ERR_HIP, "no HIP device available");
    if (h.device < 0 || h.device >= ndev) throw Error(HMX_ERR_ARG, "invalid device ordinal");
    if (h.m >= (int64_t(1) << 31) || h.n + h.R + 64 >= (int64_t(1) << 31))
        throw Error(HMX_ERR_UNSUPPORTED, "matrix too large for 32-bit device indices");
    if (h.src.kernel_id != HMX_KERNEL_CALLBACK && (h.target_dims != 1 || h.source_dims != 1))
        throw Error(HMX_ERR_UNSUPPORTED, "built-in kernels need dims == 1");
    DeviceGuard guard(h.device);
    auto op = std::make_unique<DeviceOp>();
    op->storage = storage;
    op->device = h.device;
    op->m = h.m;
    op->n = h.n;
    op->R = h.R;
    const int TK = tk_of(storage);
    const int64_t CH = 2048;  // far-field column chunk (multiple of every TK)

    // -- far field: group low-rank blocks by source cluster -----------------------
    // Per source cluster the rank rows of its blocks are stacked (T rows, in
    // this order) and cut into 64-row groups; per column chunk the row-group
    // panels are stored back to back and dispatched kWaves groups at a time.
    std::vector<Item> itemsA;     // one per (row group, chunk): generation only
    std::vector<GenA> gensA;
    std::vector<FarItem> itemsF;  // dispatch
    std::vector<ReduceGroup> groups;
    std::vector<int64_t> vrow;  // per T row: offset of its V row in the V payload
    std::vector<int64_t> trow_of_block(h.lowrank.size());
    int64_t data_off = 0, npartial = 0;
    {
        std::map<int64_t, std::vector<int64_t>> by_src;
        for (size_t b = 0; b < h.lowrank.size(); b++)
            if (h.lowrank[b].k > 0) by_src[h.lowrank[b].snode].push_back(static_cast<int64_t>(b));
        int64_t t = 0;
        for (auto& kv : by_src) {
            const auto& blocks = kv.second;
            const LowRankBlock& f = h.lowrank[blocks[0]];
            const int64_t ncols = f.c1 - f.c0 + 1;
            const int64_t tstart = t;
            for (int64_t b : blocks) {
                const LowRankBlock& L = h.lowrank[b];
                trow_of_block[b] = t;
                for (int64_t j = 0; j < L.k; j++) vrow.push_back(L.voff + j * ncols);
                t += L.k;
            }
            const int64_t Rs = t - tstart;
            const int64_t nrg = (Rs + 63) / 64;
            const int64_t nch = (ncols + CH - 1) / CH;
            int64_t pb = -1;
            if (nch > 1) {
                pb = npartial;
                npartial += Rs * nch;
                groups.push_back({pb, tstart, static_cast<int32_t>(Rs), static_cast<int32_t>(nch)});
            }
            for (int64_t c = 0; c < nch; c++) {
                const int64_t nc = std::min<int64_t>(CH, ncols - c * CH);
                const int64_t kt = (nc + TK - 1) / TK;
                for (int64_t g0 = 0; g0 < nrg; g0 += kWaves) {
                    const int64_t g1 = std::min<int64_t>(nrg, g0 + kWaves);
                    FarItem fi;
                    fi.data_off = data_off;
                    fi.col_off = (f.c0 - 1) + c * CH;
                    fi.out_off = nch > 1 ? pb + c * Rs + 64 * g0 : tstart + 64 * g0;
                    fi.ktiles = static_cast<int32_t>(kt);
                    fi.ncols = static_cast<int32_t>(nc);
                    fi.nrg = static_cast<int16_t>(g1 - g0);
                    fi.to_T = nch > 1 ? 0 : 1;
                    for (int64_t g = g0; g < g1; g++) {
                        const int32_t rows = static_cast<int32_t>(std::min<int64_t>(64, Rs - 64 * g));
                        Item it{};
                        it.data_off = data_off;
                        it.rows = rows;
                        it.ktiles = static_cast<int32_t>(kt);
                        it.ncols = static_cast<int32_t>(nc);
                        itemsA.push_back(it);
                        gensA.push_back({tstart + 64 * g, c * CH});
                        data_off += kt * TK * rows;
                        fi.rows_last = rows;
                    }
                    itemsF.push_back(fi);
                }
            }
        }
    }
    // -- leaf panels -----------------------------------------------------------------
    const std::vector<int64_t>& ls = h.target_leaf_start;  // 0-based, with end sentinel
    const int64_t nleaf = static_cast<int64_t>(ls.size()) - 1;
    std::vector<std::vector<int32_t>> dense_of(nleaf), lr_of(nleaf);
    auto leaf_range = [&](int64_t r0, int64_t r1) {  // 0-based [r0, r1)
        int64_t a = std::upper_bound(ls.begin(), ls.end(), r0) - ls.begin() - 1;
        int64_t b = std::lower_bound(ls.begin(), ls.end(), r1) - ls.begin();
        return std::make_pair(a, b);
    };
    for (size_t d = 0; d < h.dense.size(); d++) {
        auto [a, b] = leaf_range(h.dense[d].r0 - 1, h.dense[d].r1);
        for (int64_t l = a; l < b; l++) dense_of[l].push_back(static_cast<int32_t>(d));
    }
    for (size_t d = 0; d < h.lowrank.size(); d++) {
        if (h.lowrank[d].k == 0) continue;
        auto [a, b] = leaf_range(h.lowrank[d].r0 - 1, h.lowrank[d].r1);
        for (int64_t l = a; l < b; l++) lr_of[l].push_back(static_cast<int32_t>(d));
    }
    const int64_t zero_row = h.n + h.R;
    std::vector<Item> itemsB;
    std::vector<int32_t> colidx;
    std::vector<ColDesc> cdesc;
    const bool builtin = h.src.kernel_id != HMX_KERNEL_CALLBACK;
    for (int64_t l = 0; l < nleaf; l++) {
        for (int64_t r0 = ls[l]; r0 < ls[l + 1]; r0 += 64) {
            const int32_t rows = static_cast<int32_t>(std::min<int64_t>(64, ls[l + 1] - r0));
            const int64_t col_off = static_cast<int64_t>(colidx.size());
            for (int32_t d : dense_of[l]) {
                const DenseBlock& D = h.dense[d];
                const int64_t nn = D.c1 - D.c0 + 1;
                for (int64_t j = 0; j < nn; j++) {
                    colidx.push_back(static_cast<int32_t>(D.c0 - 1 + j));
                    if (builtin)
                        cdesc.push_back({h.smap[D.c0 - 1 + j] - 1, 0, 0});
                    else
                        cdesc.push_back({D.off + (r0 - (D.r0 - 1)) * nn + j, static_cast<int32_t>(nn), 2});
                }
            }
            for (int32_t b : lr_of[l]) {
                const LowRankBlock& L = h.lowrank[b];
                for (int64_t j = 0; j < L.k; j++) {
                    colidx.push_back(static_cast<int32_t>(h.n + trow_of_block[b] + j));
                    cdesc.push_back({L.uoff + (r0 - (L.r0 - 1)) * L.k + j, static_cast<int32_t>(L.k), 1});
                }
            }
            while ((colidx.size() - col_off) % TK) {
                colidx.push_back(static_cast<int32_t>(zero_row));
                cdesc.push_back({0, 0, 3});
            }
            const int64_t K = static_cast<int64_t>(colidx.size()) - col_off;
            Item it;
            it.data_off = data_off;
            it.col_off = col_off;
            it.out_off = r0;
            it.rows = rows;
            it.ktiles = static_cast<int32_t>(K / TK);
            it.ncols = static_cast<int32_t>(K);
            it.to_T = 0;
            itemsB.push_back(it);
            data_off += K * rows;
        }
    }
    op->panel_elems = data_off;
    op->npartial = npartial;

    // -- upload and generate ---------------------------------------------------------
    std::vector<int32_t> tmap0(h.m), inv(h.n);
    for (int64_t i = 0; i < h.m; i++) tmap0[i] = static_cast<int32_t>(h.tmap[i] - 1);
    for (int64_t i = 0; i < h.n; i++) inv[h.smap[i] - 1] = static_cast<int32_t>(i);
    op->tmap0 = upload(tmap0);
    op->inv_smap = upload(inv);
    HIPCHECK(hipMalloc(&op->panel, std::max<int64_t>(data_off, 1) * elem_size(storage)));
    op->colidx = upload(colidx);
    op->ncol = static_cast<int64_t>(colidx.size());

    double *dU = nullptr, *dV = nullptr, *dD = nullptr, *dxt = nullptr, *dxs = nullptr;
    ColDesc* dcd = nullptr;
    Item *dA = nullptr, *dB = nullptr;
    GenA* dgen = nullptr;
    int64_t* dvrow = nullptr;
    try {
        dU = upload(h.U);
        dV = upload(h.V);
        if (!builtin) {
            h.ensure_dense_payload();
            dD = upload(h.D);
        } else {
            dxt = upload(h.xt_copy);
            dxs = upload(h.xs_copy);
        }
        dcd = upload(cdesc);
        dA = upload(itemsA);
        dB = upload(itemsB);
        dgen = upload(gensA);
        dvrow = upload(vrow);
        const int kid = h.src.kernel_id;
        auto gen = [&](auto tag) {
            using TS = decltype(tag);
            TS* P = static_cast<TS*>(op->panel);
            if (!itemsB.empty())
                hipLaunchKernelGGL(k_gen_leaf<TS>, dim3(itemsB.size()), dim3(256), 0, 0, dB, P, dcd,
                                   op->tmap0, dxt, dxs, dU, dD, kid);
            if (!itemsA.empty())
                hipLaunchKernelGGL(k_gen_far<TS>, dim3(itemsA.size()), dim3(256), 0, 0, dA, dgen, P,
                                   dvrow, dV);
        };
        if (storage == HMX_F64) gen(double{});
        else if (storage == HMX_F32) gen(float{});
        else gen(bf16s{});
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipDeviceSynchronize());
    } catch (...) {
        for (void* p : {(void*)dU, (void*)dV, (void*)dD, (void*)dxt, (void*)dxs, (void*)dcd,
                        (void*)dA, (void*)dB, (void*)dgen, (void*)dvrow})
            if (p) (void)hipFree(p);
        device_destroy(op.release());
        throw;
    }
    for (void* p : {(void*)dU, (void*)dV, (void*)dD, (void*)dxt, (void*)dxs, (void*)dcd,
                    (void*)dgen, (void*)dvrow, (void*)dA, (void*)dB})
        if (p) (void)hipFree(p);

    order_far_items(itemsF);
    order_leaf_items(itemsB);
    op->itemsA = upload(itemsF);
    op->nA = static_cast<int64_t>(itemsF.size());
    op->itemsB = upload(itemsB);
    op->nB = static_cast<int64_t>(itemsB.size());
    op->groups = upload(groups);
    op->nG = static_cast<int64_t>(groups.size());
    h.dev[storage].reset(op.release());
}

// ---- product ---------------------------------------------------------------------
namespace {

int64_t s_rows(const DeviceOp& op) { return op.n + op.R + 1 + tk_of(op.storage); }

Scratch& get_scratch(DeviceOp& op, hipStream_t stream, int W) {
    std::lock_guard<std::mutex> g(op.mu);
    Scratch& s = op.scratch[static_cast<void*>(stream)];
    const int64_t srows = s_rows(op);
    size_t sel = op.storage == HMX_F64 ? 8 : (W == 16 && op.storage == HMX_BF16 ? 2 : 4);
    size_t pel = op.storage == HMX_F64 ? 8 : 4;
    void*& Sp = W == 16 ? s.S16 : s.S1;
    void*& Pp = W == 16 ? s.P16 : s.P1;
    if (!Sp) {
        HIPCHECK(hipMalloc(&Sp, srows * W * sel));
        HIPCHECK(hipMemset(Sp, 0, srows * W * sel));  // zero row and padding rows
        HIPCHECK(hipMalloc(&Pp, std::max<int64_t>(op.npartial, 1) * W * pel));
    }
    return s;
}

struct StageTimer {
    DeviceOp& op;
    hipStream_t st;
    int stage;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(DeviceOp& o, hipStream_t s, int k) : op(o), st(s), stage(k) {
        if (!op.profiling) return;
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        HIPCHECK(hipEventRecord(a, st));
    }
    ~StageTimer() {
        if (!op.profiling || !a) return;
        (void)hipEventRecord(b, st);
        std::lock_guard<std::mutex> g(op.mu);
        op.ev[stage].push_back({a, b});
    }
};

template <typename TS, int W, typename TIO>
void run(DeviceOp& op, const TIO* X, int64_t ldx, TIO* Y, int64_t ldy, int nrhs, hipStream_t st) {
    using s_t = typename SType<TS, W>::type;
    using acc_t = typename Store<TS>::acc;
    Scratch& sc = get_scratch(op, st, W);
    s_t* S = static_cast<s_t*>(W == 16 ? sc.S16 : sc.S1);
    acc_t* P = static_cast<acc_t*>(W == 16 ? sc.P16 : sc.P1);
    const TS* panel = static_cast<const TS*>(op.panel);
    const int64_t s_total = s_rows(op) * W * static_cast<int64_t>(sizeof(s_t));
    if (s_total >= (int64_t(1) << 32))
        throw Error(HMX_ERR_UNSUPPORTED, "source/T staging array exceeds 4 GiB (32-bit buffer offsets)");
    const uint32_t s_bytes = static_cast<uint32_t>(s_total);
    // streamed panel loads marked non-temporal (HMX_PANEL_NT=1) or default policy

    for (int rhs0 = 0; rhs0 < nrhs; rhs0 += W) {
        {
            StageTimer t(op, st, 0);
            hipLaunchKernelGGL((k_gather<s_t, W, TIO>), dim3((op.n + 255) / 256), dim3(256), 0, st, X,
                               ldx, op.inv_smap, op.n, S, rhs0, nrhs);
        }
        if (op.nA) {
            StageTimer t(op, st, 1);
            hipLaunchKernelGGL((k_far<TS, W, kPanelPolicy>), dim3(op.nA), dim3(kWaves * 64), 0, st,
                               op.itemsA, panel, S, op.n, s_bytes, P);
        }
        if (op.nG) {
            StageTimer t(op, st, 2);
            hipLaunchKernelGGL((k_reduce<TS, W>), dim3(op.nG), dim3(256), 0, st, op.groups, P, S,
                               op.n);
        }
        if (op.nB) {
            StageTimer t(op, st, 3);
            hipLaunchKernelGGL((k_leaf<TS, W, kPanelPolicy, TIO>), dim3(op.nB), dim3(kWaves * 64),
                               0, st, op.itemsB, panel, op.colidx, S, s_bytes, op.tmap0, Y, ldy,
                               rhs0, nrhs);
        }
    }
    HIPCHECK(hipGetLastError());
}

template <typename TS>
void run_io(DeviceOp& op, const void* X, int64_t ldx, void* Y, int64_t ldy, int nrhs, int io,
            hipStream_t st) {
    if (io == HMX_F64) {
        if (nrhs == 1)
            run<TS, 1, double>(op, (const double*)X, ldx, (double*)Y, ldy, nrhs, st);
        else
            run<TS, 16, double>(op, (const double*)X, ldx, (double*)Y, ldy, nrhs, st);
    } else {
        if (nrhs == 1)
            run<TS, 1, float>(op, (const float*)X, ldx, (float*)Y, ldy, nrhs, st);
        else
            run<TS, 16, float>(op, (const float*)X, ldx, (float*)Y, ldy, nrhs, st);
    }
}

}  // namespace

void device_matvec(HMatrix& h, int storage, const void* X, int64_t ldx, void* Y, int64_t ldy,
                   int nrhs, int io_dtype, void* stream) {
    if (storage < 0 || storage > 2 || !h.dev[storage])
        throw Error(HMX_ERR_STATE, "storage precision not materialised on the device");
    if (io_dtype != HMX_F64 && io_dtype != HMX_F32)
        throw Error(HMX_ERR_ARG, "io_dtype must be HMX_F64 or HMX_F32");
    if (nrhs <= 0) return;
    if (!X || !Y) throw Error(HMX_ERR_ARG, "null X or Y");
    if (ldx < h.n || ldy < h.m) throw Error(HMX_ERR_ARG, "leading dimension smaller than the size");
    DeviceOp& op = *h.dev[storage];
    DeviceGuard guard(op.device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (storage == HMX_F64)
        run_io<double>(op, X, ldx, Y, ldy, nrhs, io_dtype, st);
    else if (storage == HMX_F32)
        run_io<float>(op, X, ldx, Y, ldy, nrhs, io_dtype, st);
    else
        run_io<bf16s>(op, X, ldx, Y, ldy, nrhs, io_dtype, st);
}

void device_profile_enable(HMatrix& h, bool on) {
    for (auto& d : h.dev)
        if (d) d->profiling = on;
}

void device_profile_read(HMatrix& h, double* ms, int64_t* launches, bool reset) {
    for (int k = 0; k < 4; k++) {
        ms[k] = 0;
        launches[k] = 0;
    }
    for (auto& d : h.dev) {
        if (!d) continue;
        DeviceGuard guard(d->device);
        std::lock_guard<std::mutex> g(d->mu);
        for (int k = 0; k < 4; k++) {
            for (auto& pr : d->ev[k]) {
                HIPCHECK(hipEventSynchronize(pr.second));
                float t = 0;
                HIPCHECK(hipEventElapsedTime(&t, pr.first, pr.second));
                d->ms[k] += t;
                d->launches[k] += 1;
                (void)hipEventDestroy(pr.first);
                (void)hipEventDestroy(pr.second);
            }
            d->ev[k].clear();
            ms[k] += d->ms[k];
            launches[k] += d->launches[k];
            if (reset) {
                d->ms[k] = 0;
                d->launches[k] = 0;
            }
        }
    }
}

}  // namespace hmx
