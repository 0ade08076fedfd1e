# Restates test_tree.jl / test_hmatrix.jl against the ccall layer.  Needs julia
# and the built libhmx.so; the GPU part needs AMDGPU.jl and an MI355X.
using HMatrixGPU, Test, LinearAlgebra, Random
Random.seed!(10)

@testset "tree" begin
    coordinates = rand(3, 1000); coordinates[3, :] .= 0
    s = info(ClusterTree(coordinates; max_points_per_leaf=10))
    @test s[:max_points] <= 10
end

@testset "hmatrix circle" begin
    N = 1000
    pts = hcat([[sin(i * 2π / N), cos(i * 2π / N), 0] for i in 1:N]...)
    struct Circle <: AbstractMatrix{Float64}; X::Matrix{Float64}; end
    Base.size(K::Circle) = (size(K.X, 2), size(K.X, 2))
    Base.getindex(K::Circle, i::Int, j::Int) = (d = norm(K.X[:, i] .- K.X[:, j]); d < 1e-12 ? 1.0 : -0.5 / pi * log(d))
    K = Circle(pts)
    c = ClusterTree(pts; max_points_per_leaf=64)
    H = HMatrix(K, c, c; eta=1.5, eps=1e-6, flatten=false)
    @test info(HMatrix(K, c, c; eta=1.5, eps=1e-6))["compression_ratio"] > 3
end
