# ChannelFlowMI355XExt — MI355X (gfx950) device backend of ChannelFlow.jl's
# CNRK2 step, for fields stored in AMDGPU.ROCArray.
#
# A package extension in the layout the reference uses for its CUDA backend
# (Project.toml:32-38 [weakdeps]/[extensions], ext/ChannelFlowCUDAExt.jl).
# Register it in ChannelFlow.jl's Project.toml with
#
#   [weakdeps]
#   AMDGPU = "21141c5a-9bdb-4563-92ae-f87d6854732e"
#   [extensions]
#   ChannelFlowMI355XExt = "AMDGPU"
#
# and put libcfhip.so (channelflow.jl_amd/build.py) where LIB points, or set
# CHANNELFLOW_HIP_LIB.  The C ABI is include/channelflow_hip.h.
#
# What it overrides, by dispatch on ROCArray-backed spectral fields:
#   step!(scheme, nonlinear, U, P, t; forcing, pressuregradient | bulkvelocity)
#       src/timesteppers/cnrk2.jl:149-207 -> cf_step
# and what it adds beside it (methods of the reference's generic functions):
#   Adapt.adapt_structure(::Type{<:ROCArray}, ::ChannelFlowProblem)
#       ext/ChannelFlowCUDAExt.jl:18-32 (fields and factors move to the device)
#   random_state(problem, epsilon; seed)   initialization.jl:32-51 -> cf_random_state
# The diagnostics (cf_dot, cf_dissipation, ...) take a problem handle; a
# caller reaches them with handle(problem.scheme, form) and the ccall pattern
# of step! below.
#
# Where the device problem lives: CNRK2 is immutable and has no slot for a
# handle, so the extension keeps one HipProblem per (scheme, form) in a
# WeakKeyDict keyed by the scheme's first Q component's storage, a mutable
# ROCArray that lives exactly as long as the scheme.  The handle factors the
# three stage solvers once (cf_problem_create) from the scheme's grid, nu, dt
# and base flow.  The scheme stores the base flow as Chebyshev coefficients
# (cnrk2.jl:76); cf_config takes Lobatto values, so they are evaluated back
# (a round trip at FP64 roundoff).
#
# This file is not exercised in this repository (no Julia toolchain here);
# the same ABI calls are exercised by the Python host package and its tests.
module ChannelFlowMI355XExt

using AMDGPU, Adapt
import ChannelFlow as CF
import ChannelFlow: step!

const LIB = get(ENV, "CHANNELFLOW_HIP_LIB", joinpath(@__DIR__, "..", "deps", "libcfhip.so"))

# ROCArray-backed spectral field, the analogue of the reference's CuSpectral
const RocSpectral{T} = CF.SpectralField{T,A} where {T<:AbstractFloat,A<:ROCArray{Complex{T},3}}

# ---------------------------------------------------------------------------
# C ABI
# ---------------------------------------------------------------------------
struct CfConfig                        # mirrors cf_config in channelflow_hip.h
    Nx::Cint; Ny::Cint; Nz::Cint
    Lx::Cdouble; Lz::Cdouble
    nu::Cdouble; dt::Cdouble
    baseflow::Ptr{Cdouble}             # host: Ny values at the Lobatto points
    form::Cint
    constraint::Cint
    constraint_values::NTuple{2,Cdouble}
    device::Cint
    precision::Cint                    # 0 FP64, 1 mixed (FP32 transforms, FP64 solves)
end

const CF_FORM = Dict(CF.RotatingForm => Cint(0), CF.ConvectiveForm => Cint(1), CF.DivergenceForm => Cint(2))

function check(rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:cf_last_error, LIB), Cstring, ()))
    rc == 1 && throw(ArgumentError(msg))
    rc == 2 && throw(DimensionMismatch(msg))
    rc == 4 && throw(ErrorException("unsupported on the device: " * msg))
    throw(ErrorException(msg))
end

# device pointer of a field's storage, as the ABI's double*
devptr(f::CF.SpectralField) = reinterpret(Ptr{Cdouble}, pointer(parent(f)))
ptrs(U::CF.VectorField) = (devptr(U[1]), devptr(U[2]), devptr(U[3]))
hipstream() = AMDGPU.stream().stream   # the task-local HIP stream handle

const CALLBACK_T = (Ptr{Cvoid}, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                    Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cvoid})

# ---------------------------------------------------------------------------
# Device problem bound to a scheme
# ---------------------------------------------------------------------------
mutable struct HipProblem
    h::Ptr{Cvoid}
    form::Cint
    # callbacks registered with the handle and the objects they reach; kept
    # here so the GC roots them while the handle can call them
    forcing::Any
    nonlinear::Any
    F::Any                             # VectorField the forcing callback fills
    N::Any                             # VectorField a caller's nonlinear fills
    U::Any                             # velocity of the step in progress
end

function HipProblem(scheme::CF.CNRK2, form::Cint)
    g = scheme.solvers[1].grid
    Nx, Ny, Nz = g.physicalsize
    Lx, _, Lz = g.domainsize
    a = Array(scheme.baseflow)         # Chebyshev coefficients (cnrk2.jl:76)
    P = Ny - 1
    vals = [sum(a[n+1] * cospi(j * n / P) for n in 0:P) for j in 0:P]
    cfg = Ref(CfConfig(Nx, Ny, Nz, Lx, Lz, scheme.nu, scheme.dt, pointer(vals), form,
                       Cint(0), (0.0, 0.0), Cint(AMDGPU.device_id(AMDGPU.device()) - 1), Cint(0)))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve vals check(ccall((:cf_problem_create, LIB), Cint, (Ref{CfConfig}, Ref{Ptr{Cvoid}}), cfg, out))
    hp = HipProblem(out[], form, nothing, nothing, nothing, nothing, nothing)
    finalizer(hp) do x
        x.h == C_NULL || ccall((:cf_problem_destroy, LIB), Cint, (Ptr{Cvoid},), x.h)
        x.h = C_NULL
    end
    return hp
end

# one handle per (scheme, form), alive as long as the scheme's workspaces
const HANDLES = WeakKeyDict{Any,Dict{Cint,HipProblem}}()
const HANDLES_LOCK = ReentrantLock()

function handle(scheme::CF.CNRK2, form::Cint)
    key = parent(scheme.Q[1])
    lock(HANDLES_LOCK) do
        byform = get!(() -> Dict{Cint,HipProblem}(), HANDLES, key)
        get!(() -> HipProblem(scheme, form), byform, form)
    end
end

nlform(nl::CF.NonLinearTerm) = CF_FORM[typeof(nl).parameters[2]]
nlform(nl) = Cint(0)   # a caller's callable: the handle's own form is unused

# ---------------------------------------------------------------------------
# Callbacks: the library calls back once per stage, on the step's stream,
# with the stage time.  They enqueue device work and return 0, or 1 to stop
# the step with ArgumentError.  Neither may call this handle's transforms or
# nonlinear term: the stage's nonlinear term is live in its workspaces.
# ---------------------------------------------------------------------------
# forcing(tstage, U, F) (forcing.jl:3-11, cnrk2.jl:173-175): F is zeroed and
# the caller's forcing adds into it, as it adds into N in the reference
function forcing_trampoline(user::Ptr{Cvoid}, tstage::Cdouble, U0, U1, U2, F0, F1, F2, stream::Ptr{Cvoid})::Cint
    hp = unsafe_pointer_to_objref(user)::HipProblem
    try
        for c in 1:3
            fill!(parent(hp.F[c]), 0)
        end
        hp.forcing(tstage, hp.U, hp.F)
        return Cint(0)
    catch
        return Cint(1)
    end
end

# nonlinear(tstage, U, N) (cnrk2.jl:121-126, 172) for a callable other than
# the problem's NonLinearTerm: it overwrites N in coefficients
function nonlinear_trampoline(user::Ptr{Cvoid}, tstage::Cdouble, U0, U1, U2, N0, N1, N2, stream::Ptr{Cvoid})::Cint
    hp = unsafe_pointer_to_objref(user)::HipProblem
    try
        hp.nonlinear(tstage, hp.U, hp.N)
        return Cint(0)
    catch
        return Cint(1)
    end
end

function bind_forcing!(hp::HipProblem, forcing, U)
    forcing === hp.forcing && return
    if forcing isa CF.NoForcing
        check(ccall((:cf_set_forcing, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                    hp.h, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL))
        hp.forcing, hp.F = forcing, nothing
        return
    end
    hp.F === nothing && (hp.F = similar(U))
    fn = @cfunction(forcing_trampoline, Cint, $CALLBACK_T)
    check(ccall((:cf_set_forcing, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                hp.h, fn, pointer_from_objref(hp), ptrs(hp.F)...))
    hp.forcing = forcing
end

function bind_nonlinear!(hp::HipProblem, nl, U)
    nl === hp.nonlinear && return
    if nl isa CF.NonLinearTerm
        check(ccall((:cf_set_nonlinear, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                    hp.h, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL))
        hp.nonlinear, hp.N = nl, nothing
        return
    end
    hp.N === nothing && (hp.N = similar(U))
    fn = @cfunction(nonlinear_trampoline, Cint, $CALLBACK_T)
    check(ccall((:cf_set_nonlinear, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                hp.h, fn, pointer_from_objref(hp), ptrs(hp.N)...))
    hp.nonlinear = nl
end

# ---------------------------------------------------------------------------
# step! (src/timesteppers/cnrk2.jl:149-207) for ROCArray-backed fields.  Same
# signature and the same checks before any work is enqueued; the three stages
# run on the device on the task's HIP stream, asynchronously (as CUDA.jl
# calls are in the reference's GPU path, docs/src/gpu.md:29-31).
# ---------------------------------------------------------------------------
function step!(scheme::CF.CNRK2{S,F},
               nonlinear,
               U::CF.VectorField{F},
               P::F,
               t::Real;
               forcing=CF.NoForcing(),
               pressuregradient::Union{Nothing,NTuple{2,Real}}=nothing,
               bulkvelocity::Union{Nothing,NTuple{2,Real}}=nothing) where {S, F<:RocSpectral{Float64}}
    isnothing(pressuregradient) || isnothing(bulkvelocity) ||
        throw(ArgumentError("specify either pressuregradient or bulkvelocity"))
    g = scheme.solvers[1].grid
    for field in (U.components..., P)
        CF.grid(field) == g || throw(ArgumentError("fields must use the scheme's grid"))
        size(field) == CF.spectralsize(g, CF.NotPadded()) ||
            throw(DimensionMismatch("fields must have the resolved spectral size"))
    end
    # neither keyword: zero pressure gradient, as the generic step! does
    kind, vals = isnothing(bulkvelocity) ?
        (Cint(0), isnothing(pressuregradient) ? (0.0, 0.0) : Float64.(pressuregradient)) :
        (Cint(1), Float64.(bulkvelocity))
    hp = handle(scheme, nlform(nonlinear))
    hp.U = U
    bind_nonlinear!(hp, nonlinear, U)
    bind_forcing!(hp, forcing, U)
    v = Cdouble[vals[1], vals[2]]
    tout = Ref{Cdouble}(0)
    GC.@preserve U P v begin
        check(ccall((:cf_step, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cint,
                     Ptr{Cdouble}, Ref{Cdouble}, Ptr{Cvoid}),
                    hp.h, ptrs(U)..., devptr(P), Float64(t), kind, v, tout, hipstream()))
    end
    return tout[]
end

# ---------------------------------------------------------------------------
# Beside the step
# ---------------------------------------------------------------------------
# adapt(ROCArray, problem): fields and factors move to the device; the plans
# of the CPU NonLinearTerm are not needed on this path (the device handle
# carries its own), so the term is rebuilt on the device storage as the
# reference's CUDA extension does (ext/ChannelFlowCUDAExt.jl:18-32)
function Adapt.adapt_structure(to::Type{<:ROCArray}, p::CF.ChannelFlowProblem)
    scheme = Adapt.adapt(to, p.scheme)
    U = scheme.N[1]
    u = CF.PhysicalField(AMDGPU.zeros(Float64, CF.physicalsize(p.grid, CF.Padded())), p.grid)
    form = typeof(p.nlterm).parameters[2]()
    nl = CF.NonLinearTerm(u, U, scheme.baseflow; form=form)
    return CF.ChannelFlowProblem(p.grid, p.baseflow, nl, scheme, Adapt.adapt(to, p.forcing), p.constraint)
end

# random_state(problem, epsilon; seed) (initialization.jl:32-51) on the
# device: Philox4x32-10 noise, FFT, project!, pressure, nothing on the host.
# The noise stream is this library's documented one, not Julia's Xoshiro.
function CF.random_state(p::CF.ChannelFlowProblem{G,B,NL,S}, epsilon::Real; seed::Integer=42) where
        {G, B, NL, S<:CF.CNRK2{<:Any,<:RocSpectral{Float64}}}
    hp = handle(p.scheme, nlform(p.nlterm))
    st = CF.zero_state(p.grid)
    U, P = Adapt.adapt(ROCArray, CF.velocity(st)), Adapt.adapt(ROCArray, CF.stagepressure(st))
    check(ccall((:cf_random_state, LIB), Cint,
                (Ptr{Cvoid}, Cdouble, Culonglong, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cvoid}),
                hp.h, Float64(epsilon), UInt64(seed), ptrs(U)..., devptr(P), hipstream()))
    return CF.State(U, P)
end
