# MarchBridge (AMDGPU.jl) -- drive the gfx950 march libraries built by build_mhd.sh, build_ct.sh
# and build_march.sh (or libglmmhd.so itself) from Julia over shared device memory.
#
# The caller-side API of march_bridge/MarchBridge.jl:25-81 (open_lib, close_lib, set_dtdx, set_glm,
# set_gamma, run!) with ROCArray planes in place of CuArray planes, the same assertions and the
# same errors, plus the entry points the MI355X libraries add: set_grid (runtime grid), run_f64!
# (FP64 state) and the CT march of libct<N>.so (ct_run!, ct_regs, ct_shmem; spike_ctm.cu:299-328).
#
# State: NV SoA planes, each an nx x ny x nz ROCArray (x fastest), in the order the library's
# march_nv() implies: 9 = (rho, mx, my, mz, E, Bx, By, Bz, psi); 8 = CT (rho, mx, my, mz, E, bxf,
# byf, bzf); 5 = hydro (rho, mx, my, mz, E); 7 = hydro + two rho*s scalars.  The result of run!
# is always left in q; o is scratch.
#
#     using AMDGPU; include("glmmhdturb_amd/julia/MarchBridge.jl"); using .MarchBridge
#     m = MarchBridge.open_lib("libmhd480.so")
#     q = [AMDGPU.zeros(Float32, m.N, m.N, m.N) for _ in 1:m.NV]; o = [similar(x) for x in q]
#     MarchBridge.set_gamma(m, 5/3); MarchBridge.set_dtdx(m, 0.02); MarchBridge.set_glm(m, ch, fac)
#     ms = MarchBridge.run!(m, q, o, 30)
module MarchBridge

using AMDGPU, Libdl

struct Lib
    h::Ptr{Cvoid}
    f_dt::Ptr{Cvoid}      # march_set_dtdx
    f_run::Ptr{Cvoid}     # march_run_dev (C_NULL in the 8-plane CT library, which has no such export)
    f_glm::Ptr{Cvoid}     # march_set_glm (9-plane GLM libraries only)
    f_ct::Ptr{Cvoid}      # ct_run (GLM and CT libraries)
    NV::Int
    N::Int
end

_sym(h, s) = (f = dlsym(h, s; throw_error=false); f === nothing ? C_NULL : f)

function open_lib(path::AbstractString)
    h = dlopen(path)
    NV = Int(ccall(dlsym(h, :march_nv), Cint, ()))
    N = Int(ccall(dlsym(h, :march_nx), Cint, ()))
    Lib(h, dlsym(h, :march_set_dtdx), _sym(h, :march_run_dev), _sym(h, :march_set_glm), _sym(h, :ct_run), NV, N)
end

close_lib(m::Lib) = dlclose(m.h)

"Set the march's fixed dt/dx for the following run! / ct_run! calls."
set_dtdx(m::Lib, v::Real) = ccall(m.f_dt, Cvoid, (Cfloat,), Float32(v))

"GLM-MHD only: Dedner cleaning speed `ch` and per-step psi decay `glmfac`."
function set_glm(m::Lib, ch::Real, glmfac::Real)
    m.f_glm === C_NULL && error("this lib (NV=$(m.NV)) has no march_set_glm — not a GLM-MHD build")
    ccall(m.f_glm, Cvoid, (Cfloat, Cfloat), Float32(ch), Float32(glmfac))
end

"Adiabatic index (default 1.4)."
function set_gamma(m::Lib, g::Real)
    f = _sym(m.h, :march_set_gamma)
    f === C_NULL && error("this lib has no march_set_gamma")
    ccall(f, Cvoid, (Cfloat,), Float32(g))
end

"Runtime grid of run!/run_f64!/ct_run! (the reference fixes N at compile time)."
function set_grid(m::Lib, nx::Integer, ny::Integer, nz::Integer)
    f = _sym(m.h, :march_set_grid)
    f === C_NULL && error("this lib has no march_set_grid")
    rc = ccall(f, Cint, (Cint, Cint, Cint), nx, ny, nz)
    rc == 0 || error("march_set_grid($nx, $ny, $nz) failed with code $rc")
    nothing
end

# Host vector of device addresses, one per plane (the float** / double** ABI).
_addrs(::Type{T}, planes) where {T} = [reinterpret(Ptr{T}, UInt(UInt64(pointer(p)))) for p in planes]

function _check(m::Lib, q, o, n)
    @assert length(q) == n == length(o) "expected $n SoA planes"
    sz = size(first(q))
    @assert all(size(x) == sz for x in q) && all(size(x) == sz for x in o) "planes must share one shape"
end

"""
    run!(m, q, o, nsteps) -> elapsed_ms

`nsteps` periodic marches over NV Float32 ROCArray planes; the result is left in `q`.
"""
function run!(m::Lib, q::AbstractVector{<:ROCArray{Float32}}, o::AbstractVector{<:ROCArray{Float32}}, nsteps::Integer)
    m.f_run === C_NULL && error("this lib (NV=$(m.NV)) has no march_run_dev — use ct_run!")
    _check(m, q, o, m.NV)
    AMDGPU.synchronize()
    qa = _addrs(Float32, q); oa = _addrs(Float32, o)
    GC.@preserve q o qa oa begin
        ccall(m.f_run, Cdouble, (Ptr{Ptr{Float32}}, Ptr{Ptr{Float32}}, Cint), qa, oa, Cint(nsteps))
    end
end

"FP64 twin of run! (9-plane GLM, 5- or 7-plane hydro libraries)."
function run_f64!(m::Lib, q::AbstractVector{<:ROCArray{Float64}}, o::AbstractVector{<:ROCArray{Float64}}, nsteps::Integer)
    f = _sym(m.h, :march_run_dev_f64)
    f === C_NULL && error("this lib has no march_run_dev_f64")
    _check(m, q, o, m.NV)
    AMDGPU.synchronize()
    qa = _addrs(Float64, q); oa = _addrs(Float64, o)
    GC.@preserve q o qa oa begin
        ccall(f, Cdouble, (Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}, Cint), qa, oa, Cint(nsteps))
    end
end

"""
    ct_run!(m, q, o, cc, nsteps) -> elapsed_ms

The CT march (spike_ctm.cu:305-325) over 8 planes [U0..U4, bxf, byf, bzf]; `cc` (3 planes) is
accepted for the reference's signature and not used (the cell-centred B is formed in-kernel).
"""
function ct_run!(m::Lib, q::AbstractVector{<:ROCArray{Float32}}, o::AbstractVector{<:ROCArray{Float32}},
                 cc::AbstractVector{<:ROCArray{Float32}}, nsteps::Integer)
    m.f_ct === C_NULL && error("this lib has no ct_run")
    _check(m, q, o, 8)
    @assert length(cc) == 3 "expected 3 Bcc scratch planes"
    AMDGPU.synchronize()
    qa = _addrs(Float32, q); oa = _addrs(Float32, o); ca = _addrs(Float32, cc)
    GC.@preserve q o cc qa oa ca begin
        ccall(m.f_ct, Cdouble, (Ptr{Ptr{Float32}}, Ptr{Ptr{Float32}}, Ptr{Ptr{Float32}}, Cint), qa, oa, ca, Cint(nsteps))
    end
end

"VGPRs (arch + accumulation) and LDS bytes of the CT march kernel."
ct_regs(m::Lib) = Int(ccall(dlsym(m.h, :ct_regs), Cint, ()))
ct_shmem(m::Lib) = Int(ccall(dlsym(m.h, :ct_shmem), Cint, ()))

end # module
