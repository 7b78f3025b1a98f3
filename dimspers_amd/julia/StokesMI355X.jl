# Julia binding of libdimspers_stokes.so for Persephone (DimsPers).
#
# Keeps the reference call signatures:
#   assembly_stokes_cylindric(gr, coordinates, g, ρ, η, 𝓒, εII, T, P, PhaseID, KKidx, GGidx, MMidx)
#       (src/Stokes/StokesSolver.jl:247-404)  -> (KK, GG, MM, Fb, η)
#   StokesPcCG(U, P, KK, MM, GG, Rhs, ifree)  (src/IterativeSolvers.jl:94-168) and
#   StokesPcCG_pardiso(U, P, KK, MM, GG, Rhs, ifree)  (:1-92) on the caller's operators
#       (keywords gr, coordinates, UBC name the mesh whose pattern they live on)
#   solveStokes(U, P, T, gr, Ucartesian, Upolar, g, ρ, η, 𝓒, εII, coordinates, RotationMatrices,
#               PhaseID, UBC, KKidx, GGidx, MMidx, to; solver=:mi355x)
#       (src/Stokes/StokesSolver.jl:7-54)
# Every call passes plain pointers (GC.@preserve) and turns a non-zero status into an error.
module StokesMI355X

using LinearAlgebra: istril
using SparseArrays

const LIB = get(ENV, "DIMSPERS_STOKES_LIB", joinpath(@__DIR__, "..", "libdimspers_stokes.so"))

# ---- C structs (include/dimspers_stokes.h)
mutable struct Opts
    method::Int32; inner::Int32; stop::Int32; itmax::Int32; itmin::Int32; inner_itmax::Int32
    refine_max::Int32; cheb_degree::Int32; null_mode::Int32; pc_sweeps::Int32
    rtol::Float64; inner_rtol::Float64; mixed_inner_rtol::Float64; mixed_schur_rtol::Float64; stall_ratio::Float64
    inner_warm::Int32; fine_smoother::Int32; warm_start::Int32; pressure_pc::Int32
end
Opts() = (o = Opts(zeros(Int32, 10)..., zeros(5)..., zeros(Int32, 4)...);
          ccall((:stk_default_opts, LIB), Cvoid, (Ref{Opts},), o); o)

mutable struct Report
    converged::Int32; outer_iterations::Int32; inner_iterations::Int64; refinements::Int32; fallbacks::Int32
    schur_residual::Float64; full_residual::Float64; solve_ms::Float64
    bytes::Float64; launches::Int64; launches_floor::Int64
end
Report() = Report(0, 0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0, 0)

const SOLVE_PLAIN, SOLVE_MIXED, SOLVE_LEARNED = Int32(0), Int32(1), Int32(2)
const STOP_SCHUR, STOP_FULL = Int32(0), Int32(1)
const INNER_MG, INNER_JACOBI = Int32(0), Int32(1)
const NULL_PROJECT, NULL_DEFLATE = Int32(0), Int32(1)
const PPC_LUMPED, PPC_IC = Int32(0), Int32(1)
const ASM_FP64, ASM_FP32, ASM_CARTESIAN, ASM_NO_BC = Int32(1), Int32(2), Int32(4), Int32(8)
const MAT_K, MAT_G, MAT_GT, MAT_M = Int32(0), Int32(1), Int32(2), Int32(3)
const FSE_MODE = Dict(:NoHealing => Int32(0), :FabricDestructionOnly => Int32(1),
                      :AnnealingFabricDestruction => Int32(2))
const INSTALL_ASSEMBLY, INSTALL_STRESS = Int32(1), Int32(2)

mutable struct Context
    h::Ptr{Cvoid}
    nU::Int
    nP::Int
end

check(ctx::Union{Context,Nothing}, rc) = rc == 0 ? nothing :
    error("libdimspers_stokes: ", unsafe_string(ccall((:stk_last_error, LIB), Cstring, (Ptr{Cvoid},),
                                                      ctx === nothing ? C_NULL : ctx.h)), " (", rc, ")")

# Contexts per grid, one for every (fixed dofs, element coordinates) pair used with it.
const CONTEXTS = IdDict{Any,Vector{Tuple{Vector{Int64},Any,Context}}}()

"""One device context per (`Grid`, fixed dofs `UBC.Ω`, element `coordinates`): patterns, colouring
and multigrid structure are built once.  `UBC === nothing` stands for no fixed dof (the assembly
export, which applies no boundary condition)."""
function context(gr, coordinates, UBC)
    fixed = UBC === nothing ? Int64[] : Vector{Int64}(UBC.Ω)
    entries = get!(() -> Tuple{Vector{Int64},Any,Context}[], CONTEXTS, gr)
    for (f, coords, ctx) in entries
        f == fixed && coords === coordinates && return ctx
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    e2n = Matrix{Int64}(gr.e2n)                 # 6 x nel, 1-based, column-major
    th = Matrix{Float64}(coordinates.θ); r = Matrix{Float64}(coordinates.r)
    GC.@preserve e2n th r fixed gr begin
        rc = ccall((:stk_create, LIB), Cint,
                   (Ref{Ptr{Cvoid}}, Int64, Int64, Int64, Ptr{Int64}, Ptr{Float64}, Ptr{Float64},
                    Ptr{Float64}, Ptr{Float64}, Ptr{Int64}, Int64),
                   h, size(e2n, 2), length(gr.θ), maximum(view(e2n, 1:3, :)), e2n, th, r,
                   gr.θ, gr.r, fixed, length(fixed))
    end
    check(nothing, rc)
    ctx = Context(h[], 2 * length(gr.θ), maximum(view(e2n, 1:3, :)))
    finalizer(c -> ccall((:stk_destroy, LIB), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    push!(entries, (fixed, coordinates, ctx))
    ctx
end

function set_fields!(ctx, g, ρ, η, 𝓒)
    ip = Matrix{Float64}(η.ip)
    C = [Matrix{Float64}(getfield(𝓒, f)) for f in (:η11, :η33, :η55, :η13, :η15, :η35)]
    ptrs = [pointer(c) for c in C]
    GC.@preserve ip C ptrs ρ begin
        check(ctx, ccall((:stk_set_fields, LIB), Cint,
                         (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Ptr{Float64}}),
                         ctx.h, g, ρ.ρ, ip, size(ip, 2), ptrs))
    end
end

function export_csc(ctx, which; lower=false)
    nr, nc, nz = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    check(ctx, ccall((:stk_export_csc, LIB), Cint,
                     (Ptr{Cvoid}, Int32, Int32, Ref{Int64}, Ref{Int64}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                     ctx.h, which, lower, nr, nc, nz, C_NULL, C_NULL, C_NULL))
    colptr = Vector{Int64}(undef, nc[] + 1); rowval = Vector{Int64}(undef, nz[]); nzval = Vector{Float64}(undef, nz[])
    check(ctx, ccall((:stk_export_csc, LIB), Cint,
                     (Ptr{Cvoid}, Int32, Int32, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                     ctx.h, which, lower, C_NULL, C_NULL, C_NULL, colptr, rowval, nzval))
    SparseMatrixCSC(nr[], nc[], colptr, rowval, nzval)
end

"""GPU `assembly_stokes_cylindric`: lower-triangular KK and MM, GG, Fb in the Cartesian frame, with the
reference's positional signature (the boundary conditions are applied later, so `UBC` is optional)."""
function assembly_stokes_cylindric(gr, coordinates, g, ρ, η, 𝓒, εII, T, P, PhaseID, KKidx, GGidx, MMidx;
                                   UBC=nothing)
    ctx = context(gr, coordinates, UBC)
    set_fields!(ctx, g, ρ, η, 𝓒)
    check(ctx, ccall((:stk_assemble, LIB), Cint, (Ptr{Cvoid}, Int32), ctx.h, ASM_FP64 | ASM_CARTESIAN | ASM_NO_BC))
    Fb = Vector{Float64}(undef, ctx.nU)
    check(ctx, ccall((:stk_export_rhs, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.h, Fb))
    return export_csc(ctx, MAT_K; lower=true), export_csc(ctx, MAT_G), export_csc(ctx, MAT_M; lower=true), Fb, η
end

"""GPU `solveStokes`: assemble (rotated, with UBC's fixed dofs), Schur PCG, net-rotation removal.
`solver` picks the method of the reference's dispatch (StokesSolver.jl:41-46) or a fast one:
`:suitesparse` = StokesPcCG (FP64, zero start, lumped MM, Schur rtol 1e-8); `:pardiso` =
StokesPcCG_pardiso (FP64, starts from the incoming P, incomplete Cholesky of MM, Schur rtol 1e-6);
`:mixed` / `:learned` = FP32 Schur PCG under FP64 refinement to a 1e-8 full residual (`:learned`
needs `train!` first).  A solver name it does not know is an error."""
function solveStokes(U, P, T, gr, Ucartesian, Upolar, g, ρ, η, 𝓒, εII, coordinates, RotationMatrices,
                     PhaseID, UBC, KKidx, GGidx, MMidx, to; solver=:pardiso, rtol=nothing)
    ctx = context(gr, coordinates, UBC)
    set_fields!(ctx, g, ρ, η, 𝓒)
    check(ctx, ccall((:stk_assemble, LIB), Cint, (Ptr{Cvoid}, Int32), ctx.h, ASM_FP64 | ASM_FP32))
    o = Opts()
    if solver === :suitesparse
        o.method = SOLVE_PLAIN; o.stop = STOP_SCHUR; o.rtol = something(rtol, 1e-8)
    elseif solver === :pardiso
        o.method = SOLVE_PLAIN; o.stop = STOP_SCHUR; o.rtol = something(rtol, 1e-6)
        o.warm_start = 1; o.pressure_pc = PPC_IC
        Pin = Vector{Float64}(P)
        GC.@preserve Pin begin
            check(ctx, ccall((:stk_set_solution, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                             ctx.h, C_NULL, Pin))
        end
    elseif solver === :mixed || solver === :learned
        o.method = solver === :mixed ? SOLVE_MIXED : SOLVE_LEARNED
        o.stop = STOP_FULL; o.rtol = something(rtol, 1e-8)
    else
        error("solveStokes: unknown solver ", solver)
    end
    rep = Report()
    check(ctx, ccall((:stk_solve, LIB), Cint, (Ptr{Cvoid}, Ref{Opts}, Ref{Report}), ctx.h, o, rep))
    Ucart = similar(U); ω = Ref(0.0)
    check(ctx, ccall((:stk_updatevelocity, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
                     ctx.h, U, Ucart, ω))
    check(ctx, ccall((:stk_get_solution, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ctx.h, C_NULL, P))
    @inbounds for i in 1:2:length(U)
        Upolar[div(i, 2) + 1].x = U[i];     Upolar[div(i, 2) + 1].z = U[i + 1]
        Ucartesian[div(i, 2) + 1].x = Ucart[i]; Ucartesian[div(i, 2) + 1].z = Ucart[i + 1]
    end
    return Ucartesian, Upolar, U, Ucart, P, to
end

function _pcg_on_operators(U, P, KK, MM, GG, Rhs, ifree, gr, coordinates, UBC; warm, ppc, rtol)
    ctx = context(gr, coordinates, UBC)
    nU = length(Rhs)
    setdiff(1:nU, ifree) == sort(Vector{Int64}(UBC.Ω)) ||
        error("StokesPcCG: ifree is not the complement of UBC.Ω")
    K = SparseMatrixCSC{Float64,Int64}(KK); M = SparseMatrixCSC{Float64,Int64}(MM)
    G = SparseMatrixCSC{Float64,Int64}(GG); F = Vector{Float64}(Rhs)
    GC.@preserve K M G F begin
        check(ctx, ccall((:stk_set_operators, LIB), Cint,
                         (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Int32, Ptr{Int64}, Ptr{Int64},
                          Ptr{Float64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Int32),
                         ctx.h, Int32(istril(K)), K.colptr, K.rowval, K.nzval, Int32(istril(M)), M.colptr,
                         M.rowval, M.nzval, G.colptr, G.rowval, G.nzval, F, ASM_FP64))
    end
    if warm
        Pin = Vector{Float64}(P)
        GC.@preserve Pin begin
            check(ctx, ccall((:stk_set_solution, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                             ctx.h, C_NULL, Pin))
        end
    end
    o = Opts()
    o.method = SOLVE_PLAIN; o.inner = INNER_JACOBI; o.inner_itmax = 100000; o.stop = STOP_SCHUR; o.rtol = rtol
    o.warm_start = warm ? 1 : 0; o.pressure_pc = ppc
    # with only radial dofs fixed, K[ifree, ifree] has the near-null rotation: deflating it gives
    # the LU answer of the reference's inner solve; otherwise there is nothing to deflate
    o.null_mode = all(iseven, UBC.Ω) ? NULL_DEFLATE : NULL_PROJECT
    rep = Report()
    check(ctx, ccall((:stk_solve, LIB), Cint, (Ptr{Cvoid}, Ref{Opts}, Ref{Report}), ctx.h, o, rep))
    Uo = Vector{Float64}(undef, nU); Po = Vector{Float64}(undef, length(P))
    check(ctx, ccall((:stk_get_solution, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), ctx.h, Uo, Po))
    copyto!(U, Uo); copyto!(P, Po)
    return U, P
end

"""GPU `StokesPcCG(U, P, KK, MM, GG, Rhs, ifree)` (src/IterativeSolvers.jl:94-168) on the caller's
operators: U and P start at zero, lumped MM preconditioner, Schur rtol 1e-8, point-Jacobi CG inner
solves on K[ifree, ifree].  `gr`, `coordinates` and `UBC` name the mesh the operators live on."""
StokesPcCG(U, P, KK, MM, GG, Rhs, ifree; gr, coordinates, UBC) =
    _pcg_on_operators(U, P, KK, MM, GG, Rhs, ifree, gr, coordinates, UBC; warm=false, ppc=PPC_LUMPED, rtol=1e-8)

"""GPU `StokesPcCG_pardiso(U, P, KK, MM, GG, Rhs, ifree)` (src/IterativeSolvers.jl:1-92): keeps the
incoming P (Rhs .+= GG*P), incomplete Cholesky of MM, Schur rtol 1e-6."""
StokesPcCG_pardiso(U, P, KK, MM, GG, Rhs, ifree; gr, coordinates, UBC) =
    _pcg_on_operators(U, P, KK, MM, GG, Rhs, ifree, gr, coordinates, UBC; warm=true, ppc=PPC_IC, rtol=1e-6)

"""GPU `stress!` (src/Stress/StressCalculation.jl:244-249): fills ε, εII, τ, τII (nel × nip,
nip = size(ε.xx, 2), 7 in MAIN.jl:94,164) from the Cartesian velocity `U` (MAIN passes Ucart).
`anisotropic` selects the `Val{Anisotropic}` method (:129-242), else `Val{Isotropic}` (:23-127)."""
function stress!(ε, εII, τ, τII, η, 𝓒, U, nel, DoF_U, coordinates, SF_Stress; gr, UBC, anisotropic=true)
    ctx = context(gr, coordinates, UBC)
    nip = size(ε.xx, 2)
    ip = Matrix{Float64}(η.ip)
    C = [Matrix{Float64}(getfield(𝓒, f)) for f in (:η11, :η33, :η55, :η13, :η15, :η35)]
    ptrs = [pointer(c) for c in C]
    GC.@preserve ip C ptrs begin
        check(ctx, ccall((:stk_set_stress_fields, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Ptr{Float64}}),
                         ctx.h, ip, nip, anisotropic ? ptrs : C_NULL))
    end
    outs = (ε.xx, ε.zz, ε.xz, εII, τ.xx, τ.zz, τ.xz, τII)
    optr = [pointer(o) for o in outs]
    Uc = Vector{Float64}(U)
    GC.@preserve outs optr Uc begin
        check(ctx, ccall((:stk_stress, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Ptr{Float64}}), ctx.h, Uc, optr))
    end
    return nothing
end

"""GPU `deformation_gradient!` (StressCalculation.jl:342-349): one RK4 step of every F[iel, ip]
(a `Matrix{SMatrix{2,2,Float64,4}}`, 4 doubles per point).  Uses the stress rule set by the
preceding `stress!` call, as MAIN.jl:241-254 orders them."""
function deformation_gradient!(F, U, nel, DoF_U, coordinates, SF_Stress, Δt; gr, UBC)
    ctx = context(gr, coordinates, UBC)
    Uc = Vector{Float64}(U)
    GC.@preserve F Uc begin
        check(ctx, ccall((:stk_deformation_gradient, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Float64, Ptr{Float64}),
                         ctx.h, Uc, Δt, Ptr{Float64}(pointer(F))))
    end
    return F
end

"""GPU `solveDiffusion_threaded` (src/ThermalDiffusion/AssemblerThermal.jl:1-62): one
diffusion step with the element matrices of `assemble_element!`, `_prepare_matrices!` and
`_apply_bcs!`; the free block is solved on the device (stopping like Krylov.gmres' default)."""
function solveDiffusion_threaded(color_list, CMidx, KT, MT, FT, DoF_T, coordinates, VarT, ScratchDifussion,
                                 SF_Diffusion, valA, ρ, Δt, T, T0, TBC, to; gr, UBC)
    ctx = context(gr, coordinates, UBC)
    copyto!(T0, T)
    fixed = Vector{Int64}(TBC.Ω); vfix = Vector{Float64}(TBC.vfix); rho = Vector{Float64}(ρ)
    rep = Report()
    GC.@preserve T fixed vfix rho begin
        check(ctx, ccall((:stk_thermal_solve, LIB), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Float64, Ptr{Int64},
                          Ptr{Float64}, Int64, Float64, Ref{Report}),
                         ctx.h, T, rho, VarT.K, VarT.Cp, VarT.dQdT, Δt, fixed, vfix, length(fixed), 0.0, rep))
    end
    ΔT = T .- T0
    return T, T0, ΔT, to
end

"""GPU statistics of `write_stats` (src/IO/IOmanager.jl:277-322); appends the same line to
`stats_file` and returns (Urms, ⟨T⟩, Nu_top, Nu_bot).  `ScratchNu` is not needed: the node
lists are built once in the context."""
function write_stats(U, T, np, gr, Time, ScratchNu, stats_file; coordinates, UBC)
    ctx = context(gr, coordinates, UBC)
    out = zeros(4)
    Uc = Vector{Float64}(U); Tc = Vector{Float64}(T)
    GC.@preserve Uc Tc out begin
        check(ctx, ccall((:stk_stats, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                         ctx.h, Uc, Tc, out))
    end
    open(stats_file, "a") do io
        println(io, Float32(Time), " ", Float32(out[1]), " ", Float32(out[2]), " ", Float32(out[3]), " ",
                Float32(out[4]), " ", np)
    end
    return Tuple(out)
end

"""GPU `getFSE(F, FSE)` (src/Stress/FSE.jl:131-355): fills `FSE.fse` (a `Matrix{FiniteStrainEllipsoid}`,
6 doubles per point) from `F` (a `Matrix{SMatrix{2,2}}`) and applies the healing rule of FSE's type
parameter to F in place.  The ellipses also stay on the device for `anisotropic_tensor`."""
function getFSE(F, FSE; gr, coordinates, UBC)
    ctx = context(gr, coordinates, UBC)
    mode = get(FSE_MODE, nameof(typeof(FSE).parameters[1])) do
        error("getFSE: healing type ", typeof(FSE).parameters[1], " has no device version")
    end
    GC.@preserve F FSE begin
        check(ctx, ccall((:stk_fse, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int32, Int32, Float64, Ptr{Float64}),
                         ctx.h, Ptr{Float64}(pointer(F)), size(F, 2), mode, FSE.destruction.ϵ,
                         Ptr{Float64}(pointer(FSE.fse))))
    end
    return FSE, F
end

"""GPU `anisotropic_tensor(FSE.fse, D, Val(η))` (src/Rheology/anisotropy.jl:140-285): returns the six
nel × 7 matrices of the `StiffnessTensor` (wrap them with `StiffnessTensor(C...)`).  `fabric_shape` is
Fabrics.jl's function: evaluated at every point (exact), or at `nodes` (increasing a1 values) and
interpolated linearly on the device.  `install` (INSTALL_*) also writes 𝓒 into the device fields of the
next `solveStokes` / `stress!` of this grid."""
function anisotropic_tensor(fse, D, ::Val; gr, coordinates, UBC, fabric_shape, nodes=nothing, install=Int32(0))
    ctx = context(gr, coordinates, UBC)
    nel, nip = size(fse)
    tensor = Matrix{Float64}(D.𝓒); axes = Matrix{Float64}(D.a)
    GC.@preserve tensor axes begin
        check(ctx, ccall((:stk_set_dem, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64),
                         ctx.h, tensor, axes, size(tensor, 1), D.w))
    end
    at = nodes === nothing ? [f.a1 for f in fse] : Vector{Float64}(nodes)
    r = [fabric_shape(a, 1.0, D.parameterization) for a in at]
    r1 = Float64[x[1] for x in r]; r2 = Float64[x[2] for x in r]
    a1p = nodes === nothing ? Ptr{Float64}(C_NULL) : pointer(at)
    C = [Matrix{Float64}(undef, nel, 7) for _ in 1:6]
    ptrs = [pointer(c) for c in C]
    GC.@preserve fse at r1 r2 C ptrs begin
        check(ctx, ccall((:stk_anisotropic_tensor, LIB), Cint,
                         (Ptr{Cvoid}, Ptr{Float64}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Int32,
                          Ptr{Ptr{Float64}}),
                         ctx.h, Ptr{Float64}(pointer(fse)), nip, a1p, r1, r2, length(r1), install, ptrs))
    end
    return C
end

end # module
