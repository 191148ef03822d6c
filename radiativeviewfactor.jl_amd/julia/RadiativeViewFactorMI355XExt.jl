# RadiativeViewFactorMI355XExt.jl
# ---------------------------------------------------------------------------
# Julia glue for librvf_hip.so: registers an MI355X backend behind the
# reference's GPU hook registry (src/Assembly.jl:234-260) so that
#
#     compute_view_factors(mesh; nquad, obstruction_groups,
#                          backend = MI355XBackend(precision=:mixed))
#
# returns a ViewFactorResult computed by the HIP kernels.  The hook is
# chained, not replaced: other backends still reach the previous target.
# Knobs the reference hook drops (precision, Duffy, MC seed, device, soup
# reading) travel on the backend struct (src/Assembly.jl:77-78,243-252).
# Curve meshes (mesh_dim 1), which compute_view_factors refuses on any GPU
# backend (src/Assembly.jl:83-86), go through compute_view_factors_mi355x.
# ---------------------------------------------------------------------------
module RadiativeViewFactorMI355XExt

using RadiativeViewFactor
import RadiativeViewFactor.Assembly: _gpu_array_type, _gpu_float_type, register_gpu_hook!, _GPU_HOOK_REF
import RadiativeViewFactor.Results: ViewFactorResult
import RadiativeViewFactor.MeshIO: MeshData

const LIB = get(ENV, "RVF_MI355X_LIB", joinpath(@__DIR__, "..", "librvf_hip.so"))

Base.@kwdef struct MI355XBackend
    precision::Symbol = :fp64        # :fp64 (reference arithmetic) | :mixed
    use_duffy::Bool   = false
    seed::UInt64      = 0x0
    device::Int       = 0
    parity::Symbol    = :reference   # :reference (Q1 soup reading) | :geometric
    self_vf::Bool     = false        # the pairs (i, i) too (src/Assembly.jl:173,193); the hook drops the keyword
end

_gpu_array_type(::MI355XBackend) = Array
_gpu_float_type(::MI355XBackend) = Float64

# rvf_opts (include/rvf.h)
struct RvfOpts
    nquad::Int32
    method::Int32
    precision::Int32
    n_samples::Int32
    seed::UInt64
    flags::Int32
    reserved::Int32
end

_check(rc) = rc == 0 || error("librvf_hip: ", unsafe_string(ccall((:rvf_last_error, LIB), Cstring, ())))

const _CTX = Ref{Ptr{Cvoid}}(C_NULL)
function _ctx(device::Int)
    if _CTX[] == C_NULL
        _check(ccall((:rvf_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, _CTX))
    end
    return _CTX[]
end

const _FAM = Dict(:quad => Int8(0), :tri => Int8(1), :quad4 => Int8(2), :tri3 => Int8(3),
                  :line3 => Int8(4), :line2 => Int8(5))

function rocm_solve(mesh::MeshData, nquad::Int, b::MI355XBackend;
                    obstruction_groups::Vector{Int} = Int[], verbose::Bool = true,
                    monte_carlo::Bool = false, n_samples::Int = 10000)
    ctx = _ctx(b.device)
    elems = mesh.surface_elems
    N = length(elems)
    curve = mesh.mesh_dim == 1
    # build_gpu_arrays layout (src/GPUKernels.jl:344-385); curve meshes: Line3, Line2
    nfam = (8, 6, 4, 3, 3, 2)
    lists = [Vector{Int32}[] for _ in 1:6]
    fam = Vector{Int8}(undef, N); idx = Vector{Int32}(undef, N); grp = Vector{Int32}(undef, N)
    for (i, el) in enumerate(elems)
        f = get(_FAM, el.family) do
            error("GPU backend does not support element family $(el.family).")
        end
        (f >= 4) == curve || error("element family $(el.family) in a mesh of mesh_dim $(mesh.mesh_dim)")
        push!(lists[f + 1], Int32.(el.nodes))
        fam[i] = f; idx[i] = Int32(length(lists[f + 1])); grp[i] = Int32(el.group)
    end
    mats = [isempty(l) ? zeros(Int32, nfam[k], 0) : reduce(hcat, l) for (k, l) in enumerate(lists)]
    counts = Int64[size(m, 2) for m in mats]
    hmesh = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve mats begin
        if curve
            _check(ccall((:rvf_curve_mesh_upload, LIB), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64, Ptr{Int8}, Ptr{Int32}, Ptr{Int32},
                          Ptr{Int32}, Int64, Ptr{Int32}, Int64, Ref{Ptr{Cvoid}}),
                         ctx, mesh.coords, size(mesh.coords, 2), N, fam, idx, grp, mats[5], counts[5],
                         mats[6], counts[6], hmesh))
        else
            ptrs = [pointer(m) for m in mats[1:4]]
            _check(ccall((:rvf_mesh_upload, LIB), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64, Ptr{Int8}, Ptr{Int32}, Ptr{Int32},
                          Ptr{Ptr{Int32}}, Ptr{Int64}, Ref{Ptr{Cvoid}}),
                         ctx, mesh.coords, size(mesh.coords, 2), N, fam, idx, grp, ptrs, counts[1:4], hmesh))
        end
    end
    hobs = Ref{Ptr{Cvoid}}(C_NULL)
    tags = sort([g for g in unique(obstruction_groups) if haskey(mesh.group_tri_soup, g)])
    if !isempty(tags)
        soup = cat([mesh.group_tri_soup[g] for g in tags]...; dims = 3)
        tg = reduce(vcat, [fill(Int32(g), size(mesh.group_tri_soup[g], 3)) for g in tags])
        # segment soups (3,2,T) of curve meshes have one reading (RVF_SOUP_SEGMENTS)
        soup_read = curve ? Int32(3) : b.parity === :reference ? Int32(0) : Int32(1)
        _check(ccall((:rvf_obs_build, LIB), Cint,
                     (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Int32}, Int64, Ptr{Int32}, Int32, Int32,
                      Ref{Ptr{Cvoid}}),
                     ctx, hmesh[], soup, tg, size(soup, 3), Int32.(obstruction_groups),
                     Int32(length(obstruction_groups)), soup_read, hobs))
    elseif !isempty(obstruction_groups)
        @warn "obstruction_groups specified but no triangle geometry found for those groups; " *
              "proceeding without obstruction checking."
    end
    method = monte_carlo ? Int32(1) : (b.use_duffy ? Int32(2) : Int32(0))
    opts = RvfOpts(Int32(nquad), method, b.precision === :mixed ? Int32(1) : Int32(0), Int32(n_samples),
                   b.seed, Int32(1) | (b.self_vf ? Int32(16) : Int32(0)), Int32(0))  # RVF_WANT_F | RVF_SELF_VF
    G = Ref{Int32}(0)
    _check(ccall((:rvf_mesh_info, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ref{Int32}, Ptr{Int32}),
                 hmesh[], C_NULL, G, C_NULL))
    F = Matrix{Float64}(undef, N, N); A = Vector{Float64}(undef, N)
    Fg = Matrix{Float64}(undef, G[], G[]); Ag = Vector{Float64}(undef, G[]); gt = Vector{Int32}(undef, G[])
    _check(ccall((:rvf_compute_view_factors, LIB), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{RvfOpts}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                  Ptr{Float64}, Ptr{Int32}, Ptr{Cvoid}),
                 ctx, hmesh[], hobs[], opts, F, A, Fg, Ag, gt, C_NULL))
    hobs[] != C_NULL && ccall((:rvf_obs_free, LIB), Cvoid, (Ptr{Cvoid},), hobs[])
    ccall((:rvf_mesh_free, LIB), Cvoid, (Ptr{Cvoid},), hmesh[])
    group_tags = Int.(gt)
    group_names = [get(mesh.group_tags, t, string(t)) for t in group_tags]
    if verbose
        println("  Row-sum check (element level) — max |Σⱼ Fᵢⱼ - 1| : ", maximum(abs.(vec(sum(F, dims=2)) .- 1.0)))
        println("  Row-sum check (group level)   — max |Σⱼ Fᵢⱼ - 1| : ", maximum(abs.(vec(sum(Fg, dims=2)) .- 1.0)))
    end
    return ViewFactorResult(F, A, Fg, Ag, group_tags, group_names)
end

"""
    compute_view_factors_mi355x(mesh; nquad=4, obstruction_groups=Int[], backend=MI355XBackend(),
                                monte_carlo=false, n_samples=10000, verbose=true)

The MI355X solve for any mesh, curve meshes (mesh_dim 1) included, with the
keywords of compute_view_factors (src/Assembly.jl:64-73).
"""
function compute_view_factors_mi355x(mesh::MeshData; nquad::Int = 4, obstruction_groups::Vector{Int} = Int[],
                                     backend::MI355XBackend = MI355XBackend(), monte_carlo::Bool = false,
                                     n_samples::Int = 10000, verbose::Bool = true)
    return rocm_solve(mesh, nquad, backend; obstruction_groups, verbose, monte_carlo, n_samples)
end

function __init__()
    prev = _GPU_HOOK_REF[]
    register_gpu_hook!((m, nq, b, FT, AT; kw...) ->
        b isa MI355XBackend ? rocm_solve(m, nq, b; kw...) : prev(m, nq, b, FT, AT; kw...))
end

export MI355XBackend, compute_view_factors_mi355x

end # module
