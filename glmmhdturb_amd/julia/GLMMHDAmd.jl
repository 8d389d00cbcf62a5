# GLMMHDAmd -- glmmhd_turb.jl's step and driver API, and ct_mhd.jl's CT step, on libglmmhd.so (gfx950)
# through its context API (include/glmmhd.h), for AMDGPU.jl arrays.
#
# The functions keep the reference's names and arguments, so a driver written against
# glmmhd_turb.jl switches by replacing its step calls:
#   step_plm!(uold, unew, afield, p, dt, t; do_turb, riemann, recon, ...)   glmmhd_turb.jl:878-908
#   compute_dt(u, ct, p, be)                                                 :1973-1978
#   diagnostics(u, ct, p, be)                                                :2006-2018
#   Forcing / ou_advance!(f, p, be)                                          :2117-2161
#   run_turb(; N, t_end, amp, ...)                                           :2191-2241
#   CTMHD.step!(U, bxf, byf, bzf, bccx, bccy, bccz, Fx, Fy, Fz, ex, ey, ez, dt, dx, N)   ct_mhd.jl:174-184
# `p` is any object with the fields of the reference's Params (glmmhd_turb.jl:28-52).
# State arrays are (nx, ny, nz, 9) ROCArrays of Float64 (FP64 path) or Float32 (FP32 path): their 9
# variable planes are the library's SoA planes at stride nx*ny*nz.  Every call is ordered on the
# task-local AMDGPU stream.  Errors raise with the library's message.
#
# Differences that a caller sees: the Riemann solver and reconstruction are run-time choices
# (riemann in :llf, :llf_bound, :hll, :hlld; recon in :plm, :ppm, :ppm_spike; predictor :full or
# :oned; slope :moncen or :minmod), so step_plm! accepts every combination instead of the cube
# kernel's four; tb, half and nthreads are accepted and ignored (the march has its own tiling).
module GLMMHDAmd

using AMDGPU, Printf

const LIB = get(ENV, "GLMMHD_LIB", joinpath(@__DIR__, "..", "libglmmhd.so"))
const NDIM = 3

# layout of glm_params (include/glmmhd.h): 12 doubles, then 10 ints
struct GlmParams
    gamma::Cdouble; smallr::Cdouble; smallc::Cdouble; pfl::Cdouble; dt::Cdouble; dx::Cdouble
    ch::Cdouble; glm_fac::Cdouble; llf_dmin::Cdouble; llf_pmin::Cdouble; turb_min_rho::Cdouble; ramp::Cdouble
    predictor::Cint; recon::Cint; slope::Cint; riemann::Cint; mh_revert::Cint; face_guard::Cint
    update_floor::Cint; do_turb::Cint; zchunk::Cint; want_cfl::Cint
end

const PREDICTOR = Dict(:full => 0, :oned => 1)
const RECON = Dict(:plm => 0, :ppm => 1, :ppm_spike => 2)
const SLOPE = Dict(:moncen => 0, :minmod => 1)
const RIEMANN = Dict(:llf => 0, :llf_bound => 1, :hll => 2, :hlld => 3)
const GUARD_IFLO = 1
const NDIAG = 10

glm_error() = unsafe_string(ccall((:glm_error_string, LIB), Cstring, (Cint,), ccall((:glm_last_error, LIB), Cint, ())))
check(rc::Integer) = rc == 0 ? nothing : error("libglmmhd: " * glm_error())

dxof(p) = p.boxlen / p.N
pfloor(p) = p.smallr * p.smallc^2 / p.gamma

# ---- contexts: one per grid shape and precision, bound to the task-local stream ------------------
mutable struct Ctx
    h::Ptr{Cvoid}
    nx::Int; ny::Int; nz::Int; prec::Int
    forcing_g::Int
    function Ctx(nx, ny, nz, prec)
        h = ccall((:glm_create, LIB), Ptr{Cvoid}, (Cint, Cint, Cint, Cint), nx, ny, nz, prec)
        h == C_NULL && error("glm_create: " * glm_error())
        c = new(h, nx, ny, nz, prec, 0)
        finalizer(x -> ccall((:glm_destroy, LIB), Cvoid, (Ptr{Cvoid},), x.h), c)
        c
    end
end

const CTXS = Dict{NTuple{4,Int},Ctx}()

function context(u::ROCArray)
    nx, ny, nz = size(u, 1), size(u, 2), size(u, 3)
    c = get!(() -> Ctx(nx, ny, nz, sizeof(eltype(u))), CTXS, (nx, ny, nz, sizeof(eltype(u))))
    check(ccall((:glm_set_stream, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), c.h, AMDGPU.stream().stream))
    c
end

# device addresses of the nv variable planes of an (nx, ny, nz, nv) array
planes(u::ROCArray, nv::Integer=size(u, 4)) =
    [Ptr{Cvoid}(UInt(pointer(u)) + UInt((v - 1) * sizeof(eltype(u)) * size(u, 1) * size(u, 2) * size(u, 3))) for v in 1:nv]

# Dedner speed and psi decay of a step (glmmhd_turb.jl:884-885)
glm_ch(p, dt) = p.courant * dxof(p) / dt / NDIM * p.glm_ch_scale
glm_fac(p, ch, dt) = p.glm_cp_coef > 0 ? exp(-(ch * ch / (p.glm_cp_coef * p.boxlen * ch)) * dt) : 1.0

function params_of(p, dt, t; do_turb=false, riemann=:hll, recon=:plm, predictor=:full, slope=:moncen,
                   ch=nothing, want_cfl=false, zchunk=0)
    chv = ch === nothing ? glm_ch(p, dt) : ch
    GlmParams(p.gamma, p.smallr, p.smallc, pfloor(p), dt, dxof(p), chv, glm_fac(p, chv, dt),
              p.switch_llf_dmin, p.switch_llf_pmin, p.turb_min_rho, min(t / p.turb_T, 1.0),
              PREDICTOR[predictor], RECON[recon], SLOPE[slope], RIEMANN[riemann], 1, GUARD_IFLO, 1,
              Int(do_turb), zchunk, Int(want_cfl))
end

# ---- OU forcing on the context ---------------------------------------------------------------
mutable struct Forcing
    ctx::Ctx
    amp::Float64
    next_time::Float64
end

"Forcing(p, u; amp): the OU field (shells turb_kmin..turb_kmax on a turb_gs^3 grid) of u's context."
function Forcing(p, u::ROCArray; amp=p.turb_rms, g=hasproperty(p, :turb_gs) ? p.turb_gs : 64,
                 kmin=hasproperty(p, :turb_kmin) ? p.turb_kmin : 1.0, kmax=g ÷ 2)
    c = context(u)
    check(ccall((:glm_ou_setup, LIB), Cint,
                (Ptr{Cvoid}, Cint, Cdouble, Cdouble, Cdouble, Cint, Cdouble, Cdouble, Culonglong),
                c.h, g, kmin, kmax, p.turb_T, p.turb_Ndt, p.comp_frac, amp, UInt64(p.turb_seed)))
    c.forcing_g = g
    Forcing(c, amp, 0.0)
end

"ou_advance!(f, p, be): one OU step and a refresh of the acceleration field (glmmhd_turb.jl:2146-2161)."
function ou_advance!(f::Forcing, p, be=nothing)
    check(ccall((:glm_ou_advance, LIB), Cint, (Ptr{Cvoid},), f.ctx.h))
    f.next_time += p.turb_T / p.turb_Ndt
    nothing
end

# the kick reads the context's own field; a caller's (3, g, g, g) field (the reference's layout,
# component fastest) is copied into the library's three SoA planes first
_use_afield!(c::Ctx, ::Forcing) = nothing
_use_afield!(c::Ctx, ::Nothing) = nothing
function _use_afield!(c::Ctx, a::ROCArray)
    c.forcing_g > 0 || error("step_plm!: do_turb needs a Forcing on this grid first")
    soa = ROCArray(permutedims(a, (2, 3, 4, 1)))
    check(ccall((:glm_ou_set_afield, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), c.h, pointer(soa)))
    AMDGPU.synchronize()
    nothing
end

# ---- step API ----------------------------------------------------------------------------------
"""
    step_plm!(uold, unew, afield, p, dt, t; do_turb=false, riemann=:hll, recon=:plm, ...)

One step uold -> unew (glmmhd_turb.jl:878-908): Hancock predictor, reconstruction, Riemann
solver, GLM cleaning and psi damping, floors and, with do_turb, the OU kick.  `afield` is a
Forcing of this grid (or the reference's (3, g, g, g) field, copied in).
"""
function step_plm!(uold::ROCArray, unew::ROCArray, afield, p, dt, t; do_turb::Bool=false, nthreads::Int=192,
                   tb::Int=0, half::Bool=true, riemann::Symbol=:hll, recon::Symbol=:plm, predictor::Symbol=:full,
                   slope::Symbol=:moncen, ch=nothing, want_cfl::Bool=false)
    @assert size(uold) == size(unew) && size(uold, 4) == 9 "states must be (nx, ny, nz, 9)"
    c = context(uold)
    do_turb && _use_afield!(c, afield)
    prm = params_of(p, dt, t; do_turb, riemann, recon, predictor, slope, ch, want_cfl)
    q = planes(uold); o = planes(unew)
    GC.@preserve uold unew q o begin
        check(ccall((:glm_step, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ref{GlmParams}),
                    c.h, q, o, prm))
    end
    nothing
end

"Max over cells of sum_d(|v_d| + c_f) folded into the last step_plm!(...; want_cfl=true) of u's grid."
function cfl_max(u::ROCArray)
    out = Ref{Cdouble}(0)
    check(ccall((:glm_cfl_max, LIB), Cint, (Ptr{Cvoid}, Ref{Cdouble}), context(u).h, out))
    out[]
end

"compute_dt (glmmhd_turb.jl:1973-1978): courant dx / max_cells sum_d(|v_d| + c_f)."
function compute_dt(u::ROCArray, p)
    c = context(u)
    prm = params_of(p, 1.0, 0.0)
    out = Ref{Cdouble}(0)
    q = planes(u)
    GC.@preserve u q begin
        check(ccall((:glm_ctot_max, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ref{GlmParams}, Ref{Cdouble}),
                    c.h, q, prm, out))
    end
    p.courant * dxof(p) / out[]
end
compute_dt(u::ROCArray, ct, p, be) = compute_dt(u, p)

"diagnostics (glmmhd_turb.jl:2006-2018) plus the energy budget, FP64 compensated sums."
function diagnostics(u::ROCArray, p)
    c = context(u)
    prm = params_of(p, 1.0, 0.0)
    out = zeros(Cdouble, NDIAG)
    q = planes(u)
    GC.@preserve u q out begin
        check(ccall((:glm_diagnostics, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ref{GlmParams}, Ptr{Cdouble}),
                    c.h, q, prm, out))
    end
    mass, ke2 = out[1], out[2]
    vrms = sqrt(ke2 / mass)
    (mass=mass, vrms=vrms, mach=vrms / p.cs0, divbmax=out[3], rhomin=out[4], finite=out[5] != 0,
     ekin=out[6], emag=out[7], cross_helicity=out[8], etot=out[9], psi2=out[10])
end
diagnostics(u::ROCArray, ct, p, be) = diagnostics(u, p)

"ic_uniform (glmmhd_turb.jl:2164-2172): rho0, p0 = rho0 cs0^2/gamma, B = (b0, 0, 0)."
function ic_uniform!(u::ROCArray, p)
    c = context(u)
    args = Cdouble[p.rho0, p.rho0 * p.cs0^2 / p.gamma, p.b0, 0, 0, p.gamma]
    q = planes(u)
    GC.@preserve u q args check(ccall((:glm_ic, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Cdouble}),
                                      c.h, q, 1, args))
    nothing
end

"""
    run_turb(p; t_end, riemann=:hlld, recon=:ppm, precision=Float64, ...)

Driven turbulence (glmmhd_turb.jl:2191-2241): uniform IC, OU forcing, a CFL dt each step (from the
CFL max folded into the previous step), diagnostics every `ncontrol` steps.
"""
function run_turb(p; t_end=0.6, riemann::Symbol=:hlld, recon::Symbol=:ppm, predictor::Symbol=:full,
                  precision::Type=Float64, ncontrol::Int=200, verbose::Bool=true, nstepmax::Int=20000)
    N = p.N
    u1 = AMDGPU.zeros(precision, N, N, N, 9); u2 = similar(u1)
    ic_uniform!(u1, p)
    f = Forcing(p, u1)
    ou_advance!(f, p)
    t, step = 0.0, 0
    dt_next = compute_dt(u1, p)
    while t < t_end && step < nstepmax
        while t >= f.next_time
            ou_advance!(f, p)
        end
        dt = min(dt_next, f.next_time - t, t_end - t)
        step_plm!(u1, u2, f, p, dt, t; do_turb=true, riemann, recon, predictor, want_cfl=true)
        u1, u2 = u2, u1
        t += dt; step += 1
        dt_next = p.courant * dxof(p) / cfl_max(u1)
        if verbose && step % ncontrol == 0
            d = diagnostics(u1, p)
            @printf("  step=%d t=%.4f dt=%.2e Mach=%.3f divBmax=%.2e rhomin=%.3e finite=%s\n",
                    step, t, dt, d.mach, d.divbmax, d.rhomin, d.finite)
        end
    end
    d = diagnostics(u1, p)
    verbose && @printf("TURB N=%d: steps=%d t=%.3f Mach=%.3f divBmax=%.2e rhomin=%.3e finite=%s\n",
                       N, step, t, d.mach, d.divbmax, d.rhomin, d.finite)
    d
end

# ---- reverse step: the pullback of step_plm! (no reference counterpart) -------------------------
"""
    step_adjoint!(uold, lam1, lam0, afield, p, dt, t; kw...)

lam0 = (d unew / d uold)^T lam1 for the step step_plm!(uold, unew, afield, p, dt, t; kw...) took,
FP64 states only.  A ChainRulesCore rrule of a functional step wraps this one call:

    function ChainRulesCore.rrule(::typeof(stepped), u, afield, p, dt, t; kw...)
        unew = stepped(u, afield, p, dt, t; kw...)
        pullback(Δ) = (lam0 = similar(u); step_adjoint!(u, unthunk(Δ), lam0, afield, p, dt, t; kw...);
                       (NoTangent(), lam0, NoTangent(), NoTangent(), NoTangent(), NoTangent()))
        unew, pullback
    end
"""
function step_adjoint!(uold::ROCArray{Float64}, lam1::ROCArray{Float64}, lam0::ROCArray{Float64}, afield, p, dt, t;
                       do_turb::Bool=false, riemann::Symbol=:hll, recon::Symbol=:plm, predictor::Symbol=:full,
                       slope::Symbol=:moncen, ch=nothing)
    c = context(uold)
    prm = params_of(p, dt, t; do_turb, riemann, recon, predictor, slope, ch)
    af = C_NULL
    soa = nothing
    if do_turb
        g = c.forcing_g
        soa = ROCArray{Float64}(undef, g, g, g, 3)
        check(ccall((:glm_ou_copy_afield, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), c.h, pointer(soa)))
        af = Ptr{Cvoid}(UInt(pointer(soa)))
    end
    q = planes(uold); l1 = planes(lam1); l0 = planes(lam0)
    GC.@preserve uold lam1 lam0 soa q l1 l0 begin
        check(ccall((:glm_step_adjoint, LIB), Cint,
                    (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ref{GlmParams}, Ptr{Cvoid}),
                    c.h, q, l1, l0, prm, af))
        AMDGPU.synchronize()
    end
    nothing
end

# ---- CT: ct_mhd.jl's step! on the fused CT march -------------------------------------------------
module CTMHD
using AMDGPU
import ..GLMMHDAmd: LIB, GlmParams, check, context, planes

const GAMMA = 5.0 / 3.0
const SMALLR = 1e-6
const SMALLP = 1e-6

"""
    step!(U, bxf, byf, bzf, bccx, bccy, bccz, Fx, Fy, Fz, ex, ey, ez, dt, dx, N; gamma=GAMMA)

One CT step in place (ct_mhd.jl:174-184, same arguments): U is (N, N, N, 5), the face fields
(N, N, N) with bxf[i] = Bx on face i-1/2.  The fused march writes its result into the scratch
Fx (N, N, N, 8) and the 8 planes are copied back; the other scratch arrays are not used.
"""
function step!(U::ROCArray, bxf::ROCArray, byf::ROCArray, bzf::ROCArray, bccx, bccy, bccz,
               Fx::ROCArray, Fy, Fz, ex, ey, ez, dt, dx, N; gamma=GAMMA)
    @assert size(U) == (N, N, N, 5) && size(Fx, 4) >= 8 "U must be (N,N,N,5) and Fx (N,N,N,8)"
    c = context(U)
    # normal-only Hancock, PLM MonCen, HLL, no psi (ch = 1 and psi decay 1 satisfy the block's checks)
    prm = GlmParams(gamma, SMALLR, 1e-4, SMALLP, dt, dx, 1.0, 1.0, 0.0, 0.0, 1e-5, 0.0,
                    1, 0, 0, 2, 0, 0, 0, 0, 0, 0)
    q = vcat(planes(U, 5), [Ptr{Cvoid}(UInt(pointer(b))) for b in (bxf, byf, bzf)])
    o = planes(Fx, 8)
    GC.@preserve U bxf byf bzf Fx q o begin
        check(ccall((:glm_ct_step, LIB), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ref{GlmParams}, Cint),
                    c.h, q, o, prm, 1))
        AMDGPU.synchronize()
    end
    n3 = N^3
    copyto!(U, 1, Fx, 1, 5 * n3)
    copyto!(bxf, 1, Fx, 5 * n3 + 1, n3)
    copyto!(byf, 1, Fx, 6 * n3 + 1, n3)
    copyto!(bzf, 1, Fx, 7 * n3 + 1, n3)
    nothing
end

"max |div B| from the face fields (ct_mhd.jl:191-194)."
function divB_max(bxf::ROCArray, byf::ROCArray, bzf::ROCArray, dx)
    c = context(bxf)
    fb = [Ptr{Cvoid}(UInt(pointer(b))) for b in (bxf, byf, bzf)]
    q = vcat(fill(fb[1], 5), fb)      # only planes 6-8 (the face fields) are read
    out = Ref{Cdouble}(0)
    GC.@preserve bxf byf bzf q check(ccall((:glm_ct_divb_max, LIB), Cint,
                                           (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cdouble, Ref{Cdouble}), c.h, q, dx, out))
    out[]
end

end # module CTMHD

end # module GLMMHDAmd
