"""Step and driver API of the GLM-MHD solver on gfx950, mirroring glmmhd_turb.jl.

Names follow the Julia module: ``Params`` (glmmhd_turb.jl:28-52), ``dxof``/``pfloor`` (:54-55),
``step_plm`` (step_plm!, :878-908), ``compute_dt`` (:1973-1978), ``diagnostics`` (:2006-2018),
``ic_ot`` / ``ic_uniform`` (:2021-2036, :2164-2172), ``Forcing`` / ``ou_advance`` (:2117-2161),
``run_ot`` (:2041-2068) and ``run_turb`` (:2191-2241).

State: a torch tensor of shape (9, nz, ny, nx) on the GPU (x fastest), i.e. the memory layout
of the Julia (N,N,N,9) array.  Every numeric entry point runs in libglmmhd.so; there is no CPU
path.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace

from . import _lib
from ._lib import GLMParams, check, plane_array

NV = 9
NDIM = 3            # glmmhd_turb.jl:59 (kept at 3 for z-invariant 2-D runs, as the reference)

# Scheme presets: which reference code path a step reproduces.
SCHEMES = {
    # integrator_plm! / integrator_plm_ppm! (glmmhd_turb.jl:782-1014): full Hancock, mh revert,
    # iflo face revert, update floor
    "cube": dict(predictor="full", mh_revert=True, face_guard="iflo", update_floor=True),
    # step_multi! (glmmhd_turb.jl:1981-1992): as cube but without the face revert
    "multi": dict(predictor="full", mh_revert=True, face_guard="none", update_floor=True),
    # cpu_simd.jl (:31-101, :462-540): full Hancock, MonCen, LLF, no face revert
    "cpu_simd": dict(predictor="full", mh_revert=True, face_guard="none", update_floor=True,
                     riemann="llf", slope="moncen", recon="plm"),
    # spike_mhd.cu march (:227-302): normal-only Hancock, clamped faces, no floors
    "spike": dict(predictor="oned", mh_revert=False, face_guard="clamp", update_floor=False,
                  riemann="hll", recon="plm", slope="moncen"),
}


@dataclass
class Params:
    """glmmhd_turb.jl:28-52 plus the grid shape and forcing band of this implementation."""
    N: int = 256
    boxlen: float = 1.0
    gamma: float = 1.01
    courant: float = 0.7
    smallr: float = 1e-6
    smallc: float = 1e-4
    glm_ch_scale: float = 0.25
    glm_cp_coef: float = 0.18
    # Dedner cleaning speed: "ref" = courant*dx/dt/3*glm_ch_scale (glmmhd_turb.jl:884), "maxfast" =
    # the largest fast magnetosonic speed on the grid (max over cells and axes of c_f(B_d))
    glm_ch_mode: str = "ref"
    switch_llf_dmin: float = 0.0
    switch_llf_pmin: float = 0.0
    rho0: float = 1.0
    cs0: float = 1.0
    b0: float = 1.0
    turb_rms: float = 130.0
    turb_T: float = 0.3
    turb_Ndt: int = 100
    comp_frac: float = 0.5
    turb_min_rho: float = 1e-5
    turb_seed: int = 1
    # grid (nx, ny, nz default to N; dx = boxlen/N in every direction)
    nx: int | None = None
    ny: int | None = None
    nz: int | None = None
    # OU forcing band (reference: 1 <= |k| <= n/2 on a 64^3 grid, :2078-2091)
    turb_gs: int = 64
    turb_kmin: float = 1.0
    turb_kmax: float | None = None
    # uniform-IC field direction (reference: B = (b0, 0, 0), :2164-2172)
    b0_dir: tuple = (1.0, 0.0, 0.0)

    @property
    def shape(self):
        return (self.nx or self.N, self.ny or self.N, self.nz or self.N)


def dxof(p: Params) -> float:
    """glmmhd_turb.jl:54"""
    return p.boxlen / p.N


def pfloor(p: Params) -> float:
    """glmmhd_turb.jl:55"""
    return p.smallr * p.smallc ** 2 / p.gamma


def glm_speed(p: Params, dt: float, ch: float | None = None) -> tuple[float, float]:
    """Dedner ch and psi decay factor of a step (glmmhd_turb.jl:884-885).  A given ch (e.g. the
    max fast speed of the state, glm_ch_mode "maxfast") replaces the reference's dt rule."""
    dx = dxof(p)
    if ch is None:
        ch = p.courant * dx / dt / NDIM * p.glm_ch_scale
    fac = math.exp(-(ch * ch / (p.glm_cp_coef * p.boxlen * ch)) * dt) if p.glm_cp_coef > 0 else 1.0
    return ch, fac


def ramp_of(p: Params, t: float) -> float:
    """glmmhd_turb.jl:886"""
    return min(t / p.turb_T, 1.0)


def step_args(p: Params, dt: float, t: float, *, do_turb=False, riemann="hll", recon="plm",
              predictor=None, slope="moncen", scheme="cube", want_cfl=True, zchunk=0,
              ch: float | None = None, glm_fac: float | None = None) -> GLMParams:
    """Build the C parameter block of one step (host-side logic of step_plm!)."""
    sc = dict(SCHEMES[scheme])
    sc.setdefault("riemann", riemann)
    sc.setdefault("recon", recon)
    sc.setdefault("slope", slope)
    if scheme not in ("cpu_simd", "spike"):
        sc["riemann"], sc["recon"], sc["slope"] = riemann, recon, slope
    if predictor is not None:
        sc["predictor"] = predictor
    ch, fac_ = glm_speed(p, dt, ch)
    glm_fac = fac_ if glm_fac is None else glm_fac
    a = GLMParams()
    a.gamma, a.smallr, a.smallc, a.pfl = p.gamma, p.smallr, p.smallc, pfloor(p)
    a.dt, a.dx, a.ch, a.glm_fac = dt, dxof(p), ch, glm_fac
    a.llf_dmin, a.llf_pmin = p.switch_llf_dmin, p.switch_llf_pmin
    a.turb_min_rho, a.ramp = p.turb_min_rho, ramp_of(p, t)
    a.predictor = _lib.PREDICTOR[sc["predictor"]]
    a.recon = _lib.RECON[sc["recon"]]
    a.slope = _lib.SLOPE[sc["slope"]]
    a.riemann = _lib.RIEMANN[sc["riemann"]]
    a.mh_revert = int(sc["mh_revert"])
    a.face_guard = _lib.GUARD[sc["face_guard"]]
    a.update_floor = int(sc["update_floor"])
    a.do_turb = int(do_turb)
    a.zchunk = int(zchunk)
    a.want_cfl = int(want_cfl)
    return a


def hydro_args(dtdx: float, gamma: float = 1.4, *, smallr: float = 1e-6, smallp: float = 1e-6,
               fallback: tuple[float, float] | None = None, zchunk: int = 0, ppm: bool = False) -> GLMParams:
    """Parameters of the 5-variable hydro march (spike_25d.cu:180 constants: gamma 1.4,
    SMALLR = SMALLP = 1e-6).  fallback = (dmin, pmin) selects HLL with the LLF fallback below
    those floors (-DLLF_FALLBACK, (1e-3, 1e-3) in spike_25d.cu:339); ppm the parabolic-edge build
    (-DPPM1D, :467-483, which takes no fallback)."""
    a = GLMParams()
    _lib.load().glm_params_default(C.byref(a))
    a.gamma, a.dt, a.dx, a.smallr, a.pfl = gamma, dtdx, 1.0, smallr, smallp
    a.recon = _lib.RECON["ppm_spike" if ppm else "plm"]
    a.llf_dmin, a.llf_pmin = fallback if fallback is not None else (0.0, 0.0)
    a.zchunk, a.want_cfl, a.do_turb = int(zchunk), 0, 0
    return a


def march_args(dtdx: float, gamma: float = 1.4, *, ch: float = 1.0, glm_fac: float = 1.0, recon: str = "plm",
               zchunk: int = 0) -> GLMParams:
    """Parameters of the line march that the MarchBridge libraries run (spike_mhd.cu:52, :257-288;
    spike_ctm.cu:58): normal-only Hancock, MonCen, HLL, the GLM pair with CH and psi *= GLMFAC,
    faces clamped at SMALLR = SMALLP = 1e-6, no predictor revert, no update floor, fixed dt/dx.
    recon "ppm_spike" is the -DPPM build's parabolic-edge reconstruction (spike_mhd.cu:164-188)."""
    a = GLMParams()
    _lib.load().glm_params_default(C.byref(a))
    a.gamma, a.dt, a.dx, a.smallr, a.pfl, a.ch, a.glm_fac = gamma, dtdx, 1.0, 1e-6, 1e-6, ch, glm_fac
    a.predictor, a.recon, a.slope, a.riemann = (_lib.PREDICTOR["oned"], _lib.RECON[recon], _lib.SLOPE["moncen"],
                                                _lib.RIEMANN["hll"])
    a.mh_revert, a.face_guard, a.update_floor = 0, _lib.GUARD["clamp"], 0
    a.zchunk, a.want_cfl, a.do_turb = int(zchunk), 0, 0
    return a


def _torch():
    import torch
    return torch


class Solver:
    """One grid's device context (libglmmhd glm_ctx) bound to torch's current stream."""

    def __init__(self, p: Params, precision: str = "f64", device: str = "cuda:0"):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.GLMError("a GPU is required (libglmmhd has no CPU path)")
        self.lib = _lib.load()
        self.p = p
        self.nx, self.ny, self.nz = p.shape
        self.prec = precision
        self.dtype = torch.float64 if precision == "f64" else torch.float32
        self.device = torch.device(device)
        with torch.cuda.device(self.device):
            self.ctx = self.lib.glm_create(self.nx, self.ny, self.nz, 8 if precision == "f64" else 4)
            if not self.ctx:
                raise _lib.GLMError(self.lib.glm_error_string(self.lib.glm_last_error()).decode())
            check(self.lib.glm_set_stream(self.ctx, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        self.forcing = None
        self.zchunk = 0      # z planes per march tile for steps whose args leave it at 0 (0: the library's model)
        self.ct_zchunk = 0   # the same for the CT march

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.glm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ------------------------------------------------------------------------
    def new_state(self, nv=NV):
        torch = _torch()
        return torch.zeros((nv, self.nz, self.ny, self.nx), dtype=self.dtype, device=self.device)

    def _planes(self, u, nv=NV):
        if u.dtype != self.dtype or not u.is_contiguous() or tuple(u.shape) != (nv, self.nz, self.ny, self.nx):
            raise ValueError(f"state must be a contiguous {self.dtype} tensor of shape {(nv, self.nz, self.ny, self.nx)}")
        return plane_array([u[v].data_ptr() for v in range(nv)])

    # ---- initial conditions -------------------------------------------------------------
    def ic(self, u, kind: str, args):
        pa, keep = self._planes(u)
        arr = (C.c_double * len(args))(*args)
        check(self.lib.glm_ic(self.ctx, pa, _lib.IC[kind], arr))
        return u

    def ic_ot(self, u, gamma):
        return self.ic(u, "orszag_tang", [gamma])

    def ic_uniform(self, u, rho0, p0, b0, gamma):
        b = tuple(b0) if hasattr(b0, "__len__") else (b0, 0.0, 0.0)
        return self.ic(u, "uniform", [rho0, p0, b[0], b[1], b[2], gamma])

    def ic_gradient(self, u, gamma):
        return self.ic(u, "gradient", [gamma])

    # ---- step -------------------------------------------------------------------------
    def step(self, uold, unew, args: GLMParams):
        qa, kq = self._planes(uold)
        oa, ko = self._planes(unew)
        if args.zchunk == 0 and self.zchunk:
            args = GLMParams.from_buffer_copy(args)
            args.zchunk = self.zchunk
        check(self.lib.glm_step(self.ctx, qa, oa, C.byref(args)))

    def step_species(self, uold, unew, args: GLMParams, nsp: int = 2):
        """step() on (9 + nsp, nz, ny, nx) states whose last nsp planes are the conserved rho*X_s of
        passive species riding the mass flux (spike_mhd.cu -DNSP, :55-61, :218-220, :278-296)."""
        qa, kq = self._planes(uold, NV + nsp)
        oa, ko = self._planes(unew, NV + nsp)
        if args.zchunk == 0 and self.zchunk:
            args = GLMParams.from_buffer_copy(args)
            args.zchunk = self.zchunk
        check(self.lib.glm_step_species(self.ctx, qa, oa, C.byref(args), int(nsp)))

    def march_floor(self, uold, unew, args: GLMParams):
        """The memory floor of step() (spike_mhd.cu -DMEMFLOOR): the march's loads, staging and
        stores with no flux computation; unew receives uold unchanged.  For measurement."""
        qa, kq = self._planes(uold)
        oa, ko = self._planes(unew)
        if args.zchunk == 0 and self.zchunk:
            args = GLMParams.from_buffer_copy(args)
            args.zchunk = self.zchunk
        check(self.lib.glm_march_floor(self.ctx, qa, oa, C.byref(args)))

    def tune_zchunk(self, uold, unew, args: GLMParams, reps: int = 3, ct: bool = False, face_slope: int = 0) -> dict:
        """Time the march (ct=True: the CT march) for a set of z-chunkings of this grid and keep the
        fastest for later steps whose args leave zchunk at 0.  Every chunking gives the same bits
        (each chunk recomputes the faces at its boundary), so this changes speed only.  Writes only
        unew.  Candidates: 0 (the library's occupancy model) and ceil(nz / n) for n = 4..16 chunks."""
        torch = _torch()
        cands = [0] + sorted({-(-self.nz // n) for n in range(4, 17) if -(-self.nz // n) >= 8}, reverse=True)
        st = torch.cuda.current_stream()
        times = {}
        if ct:
            self.ct_zchunk = 0
            run = lambda a: self.ct_step(uold, unew, a, face_slope)  # noqa: E731
        else:
            self.zchunk = 0
            run = lambda a: self.step(uold, unew, a)  # noqa: E731
        for zc in cands:
            a = GLMParams.from_buffer_copy(args)
            a.zchunk, a.want_cfl = zc, 0
            run(a)                                        # untimed first launch
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(reps):
                run(a)
            e1.record(st)
            torch.cuda.synchronize()
            times[zc] = e0.elapsed_time(e1) / reps
        best = min(times, key=times.get)
        if ct:
            self.ct_zchunk = best
        else:
            self.zchunk = best
        return {"zchunk": best, "ms": times[best], "auto_ms": times[0]}

    def step_plm(self, uold, unew, p: Params, dt: float, t: float, *, do_turb=False, riemann="hll",
                 recon="plm", predictor=None, slope="moncen", scheme="cube", want_cfl=True, zchunk=0,
                 ch: float | None = None):
        """step_plm! (glmmhd_turb.jl:878-908): one step uold -> unew.  With glm_ch_mode "maxfast"
        the cleaning speed is the max fast speed of uold unless ch is given."""
        if do_turb and self.forcing is None:
            raise _lib.GLMError("do_turb requires setup_forcing()")
        if ch is None and p.glm_ch_mode == "maxfast":
            ch = self.fast_max(uold)
        elif p.glm_ch_mode not in ("ref", "maxfast"):
            raise ValueError(f"glm_ch_mode must be 'ref' or 'maxfast', not {p.glm_ch_mode!r}")
        a = step_args(p, dt, t, do_turb=do_turb, riemann=riemann, recon=recon, predictor=predictor,
                      slope=slope, scheme=scheme, want_cfl=want_cfl, zchunk=zchunk, ch=ch)
        self.step(uold, unew, a)

    def cfl_max(self) -> float:
        """Max signal speed folded into the last step (reads back one double)."""
        out = C.c_double()
        check(self.lib.glm_cfl_max(self.ctx, C.byref(out)))
        return out.value

    def ctot_max(self, u) -> float:
        pa, keep = self._planes(u)
        out = C.c_double()
        a = step_args(self.p, 1.0, 0.0)
        check(self.lib.glm_ctot_max(self.ctx, pa, C.byref(a), C.byref(out)))
        return out.value

    def fast_max(self, u) -> float:
        """Largest fast magnetosonic speed on the grid, max over cells and axes of c_f(B_d)."""
        pa, keep = self._planes(u)
        out = C.c_double()
        a = step_args(self.p, 1.0, 0.0)
        check(self.lib.glm_fast_max(self.ctx, pa, C.byref(a), C.byref(out)))
        return out.value

    def compute_dt(self, u, p: Params | None = None) -> float:
        """glmmhd_turb.jl:1973-1978"""
        p = p or self.p
        return p.courant * dxof(p) / self.ctot_max(u)

    def diagnostics(self, u, p: Params | None = None) -> dict:
        """glmmhd_turb.jl:2006-2018 plus E_kin, E_mag, cross helicity, sum E, sum psi^2 (FP64)."""
        p = p or self.p
        pa, keep = self._planes(u)
        out = (C.c_double * _lib.NDIAG)()
        a = step_args(p, 1.0, 0.0)
        check(self.lib.glm_diagnostics(self.ctx, pa, C.byref(a), out))
        d = dict(zip(_lib.DIAG_KEYS, list(out)))
        d["vrms"] = math.sqrt(d["ke2"] / d["mass"]) if d["mass"] > 0 else float("nan")
        d["mach"] = d["vrms"] / p.cs0
        d["finite"] = bool(d["finite"])
        return d

    # ---- OU forcing ---------------------------------------------------------------------
    def setup_forcing(self, p: Params | None = None, amp: float | None = None):
        """Forcing (glmmhd_turb.jl:2133-2143) on this context."""
        p = p or self.p
        kmax = p.turb_kmax if p.turb_kmax is not None else p.turb_gs // 2
        amp = p.turb_rms if amp is None else amp
        check(self.lib.glm_ou_setup(self.ctx, p.turb_gs, p.turb_kmin, float(kmax), p.turb_T, p.turb_Ndt,
                                    p.comp_frac, amp, p.turb_seed))
        self.forcing = Forcing(self, p, amp)
        return self.forcing

    def afield(self):
        """The acceleration field as a (3, g, g, g) torch view (device memory of the context)."""
        torch = _torch()
        if self.forcing is None:
            raise _lib.GLMError("afield requires setup_forcing()")
        g = self.p.turb_gs
        out = torch.empty((3, g, g, g), dtype=self.dtype, device=self.device)
        check(self.lib.glm_ou_copy_afield(self.ctx, C.c_void_p(out.data_ptr())))
        torch.cuda.synchronize(self.device)
        return out

    def _check_afield(self, field, who: str):
        """A forcing field must be a contiguous (3, g, g, g) tensor of the state dtype on this
        context's device: the library copies or reads 3 g^3 values from its address."""
        g = self.p.turb_gs
        if (getattr(field, "dtype", None) != self.dtype or tuple(field.shape) != (3, g, g, g)
                or not field.is_contiguous() or field.device != self.device):
            raise _lib.GLMError(f"{who}: the forcing field must be a contiguous {self.dtype} tensor of shape "
                                f"{(3, g, g, g)} on {self.device}")

    def set_afield(self, field):
        """Replace the context's forcing field by `field` ((3, g, g, g) tensor, state dtype)."""
        if self.forcing is None:
            raise _lib.GLMError("set_afield requires setup_forcing()")
        self._check_afield(field, "set_afield")
        check(self.lib.glm_ou_set_afield(self.ctx, C.c_void_p(field.data_ptr())))

    # ---- adjoint ----------------------------------------------------------------------------
    def step_adjoint(self, u0, lam1, lam0, args: GLMParams, afield=None, passes: bool = False):
        """lam0 = (dU1/dU0)^T lam1 for the step `args` taken from u0 (FP64); afield is the forcing
        field the step used when args.do_turb.  passes=True runs the back half of the reverse step
        as separate passes instead of the fused z-march (same results)."""
        qa, k0 = self._planes(u0)
        la, k1 = self._planes(lam1)
        oa, k2 = self._planes(lam0)
        if afield is not None:
            self._check_afield(afield, "step_adjoint")
        af = None if afield is None else C.c_void_p(afield.data_ptr())
        fn = self.lib.glm_step_adjoint_passes if passes else self.lib.glm_step_adjoint
        check(fn(self.ctx, qa, la, oa, C.byref(args), af))

    def adjoint_force_dot(self, reset: bool = True) -> float:
        """Sum over the reverse steps so far of (kick acceleration adjoint) . (kick acceleration)."""
        out = C.c_double()
        check(self.lib.glm_adjoint_force_dot(self.ctx, int(reset), C.byref(out)))
        return out.value

    # ---- constrained transport ------------------------------------------------------------
    def ct_step(self, qold, qnew, args: GLMParams, face_slope: int = 0, passes: bool = False, path: str = "march"):
        """One CT step (ct_mhd.jl step!, :174-184).  path "march": the fused z-march kernel;
        "passes" (or passes=True): the multi-pass kernel sequence; "tiles": one kernel over staged
        3-D tiles (spike_ct2.cu design).  All three give the same results to rounding."""
        if passes:
            path = "passes"
        fns = {"march": self.lib.glm_ct_step, "passes": self.lib.glm_ct_step_passes, "tiles": self.lib.glm_ct_step_tiles}
        if path not in fns:
            raise ValueError(f"ct_step path must be one of {sorted(fns)}, not {path!r}")
        qa, kq = self._planes(qold, 8)
        oa, ko = self._planes(qnew, 8)
        if args.zchunk == 0 and self.ct_zchunk and path == "march":
            args = GLMParams.from_buffer_copy(args)
            args.zchunk = self.ct_zchunk
        check(fns[path](self.ctx, qa, oa, C.byref(args), int(face_slope)))

    def ct_step_species(self, qold, qnew, args: GLMParams, nsp: int = 2, face_slope: int = 0):
        """ct_step (fused march) on (8 + nsp, nz, ny, nx) states whose last nsp planes are the
        conserved rho*X_s of passive species riding the mass flux (spike_ctm.cu -DNSP)."""
        qa, kq = self._planes(qold, 8 + nsp)
        oa, ko = self._planes(qnew, 8 + nsp)
        if args.zchunk == 0 and self.ct_zchunk:
            args = GLMParams.from_buffer_copy(args)
            args.zchunk = self.ct_zchunk
        check(self.lib.glm_ct_step_species(self.ctx, qa, oa, C.byref(args), int(face_slope), int(nsp)))

    def hydro_step(self, qold, qnew, args: GLMParams, nsp: int = 0, cma: bool = False, donor: bool = False):
        """One step of the hydro march (spike_25d.cu march, build_march.sh scheme) on
        (5 + nsp, nz, ny, nx) conserved states [rho, mx, my, mz, E, rho*s0, rho*s1]; nsp = 2 adds
        the passive scalars of -DSCALARS, cma their -DCMA flux, donor the donor-cell tile seams of
        -DDONOR; see hydro_args."""
        qa, kq = self._planes(qold, 5 + nsp)
        oa, ko = self._planes(qnew, 5 + nsp)
        if donor:
            check(self.lib.glm_hydro_step_donor(self.ctx, qa, oa, C.byref(args), int(nsp), int(cma)))
        elif nsp == 0 and not cma:
            check(self.lib.glm_hydro_step(self.ctx, qa, oa, C.byref(args)))
        else:
            check(self.lib.glm_hydro_step_scalars(self.ctx, qa, oa, C.byref(args), int(nsp), int(cma)))

    HYDRO_FORMATS = {"f16_prim": 1, "u16_prim": 2, "u16_log2_species": 3}

    def hydro_step_compact(self, qplanes, oplanes, args: GLMParams, fmt: str):
        """hydro_step on a compact storage format (spike_25d.cu -DF16STATE / -DU16STATE / -DU16SP,
        :64-95), FP32 contexts only: "f16_prim" (5 float16 planes rho, vx, vy, vz, p), "u16_prim"
        (5 16-bit planes, rho and p log2-coded, velocities linear) or "u16_log2_species" (5 float32
        conserved planes, then 2 16-bit planes of log2-coded species fractions).  qplanes, oplanes:
        sequences of contiguous (nz, ny, nx) tensors on the context's device."""
        if fmt not in self.HYDRO_FORMATS:
            raise ValueError(f"fmt must be one of {sorted(self.HYDRO_FORMATS)}, not {fmt!r}")
        code = self.HYDRO_FORMATS[fmt]
        nv = 7 if code == 3 else 5
        shape = (self.nz, self.ny, self.nx)
        for planes in (qplanes, oplanes):
            if len(planes) != nv:
                raise ValueError(f"{fmt} takes {nv} planes, not {len(planes)}")
            for v, x in enumerate(planes):
                wide = code == 3 and v < 5
                size = 4 if wide else 2
                if (tuple(x.shape) != shape or not x.is_contiguous() or x.element_size() != size
                        or x.device != self.device):
                    raise ValueError(f"{fmt}: plane {v} must be a contiguous {size}-byte tensor of shape {shape} "
                                     f"on {self.device}")
        qa, kq = plane_array([x.data_ptr() for x in qplanes])
        oa, ko = plane_array([x.data_ptr() for x in oplanes])
        check(self.lib.glm_hydro_step_compact(self.ctx, qa, oa, C.byref(args), int(code)))

    def ct_ctot_max(self, q) -> float:
        pa, keep = self._planes(q, 8)
        out = C.c_double()
        a = step_args(self.p, 1.0, 0.0)
        check(self.lib.glm_ct_ctot_max(self.ctx, pa, C.byref(a), C.byref(out)))
        return out.value

    def ct_divb_max(self, q, dx: float | None = None) -> float:
        pa, keep = self._planes(q, 8)
        out = C.c_double()
        check(self.lib.glm_ct_divb_max(self.ctx, pa, float(dx or dxof(self.p)), C.byref(out)))
        return out.value


class Forcing:
    """glmmhd_turb.jl:2117-2131: OU state lives in the library context; next_time is host-side."""

    def __init__(self, solver: Solver, p: Params, amp: float):
        self.solver, self.p, self.amp = solver, p, amp
        self.next_time = 0.0

    def advance(self):
        """ou_advance! (glmmhd_turb.jl:2146-2161)"""
        check(self.solver.lib.glm_ou_advance(self.solver.ctx))
        self.next_time += self.p.turb_T / self.p.turb_Ndt

    @property
    def nsteps(self) -> int:
        return int(self.solver.lib.glm_ou_steps(self.solver.ctx))


def ou_advance(f: Forcing) -> None:
    f.advance()


def run_ot(N: int = 128, t_end: float = 0.5, *, nz: int | None = None, riemann="hlld", recon="plm",
           predictor="full", slope="moncen", scheme="multi", courant=0.7, precision="f64", verbose=True,
           ch_mode: str = "ref"):
    """glmmhd_turb.jl:2041-2068: Orszag-Tang to t_end with a CFL dt each step, no forcing.
    ch_mode "maxfast" takes the cleaning speed from the max fast speed of each step's state."""
    p = Params(N=N, gamma=5.0 / 3.0, cs0=1.0, courant=courant, nz=nz, glm_ch_mode=ch_mode)
    s = Solver(p, precision)
    u1, u2 = s.new_state(), s.new_state()
    s.ic_ot(u1, p.gamma)
    t, step = 0.0, 0
    while t < t_end:
        dt = s.compute_dt(u1)
        dt = min(dt, t_end - t)
        s.step_plm(u1, u2, p, dt, t, riemann=riemann, recon=recon, predictor=predictor, slope=slope,
                   scheme=scheme, want_cfl=False)
        u1, u2 = u2, u1
        t += dt
        step += 1
    d = s.diagnostics(u1)
    if verbose:
        print(f"OT N={N} {riemann}: steps={step} t={t:.3f} vrms={d['vrms']:.4f} divBmax={d['divbmax']:.3e} "
              f"rhomin={d['rhomin']:.4f} finite={d['finite']}")
    return d, u1, s


def run_turb(N: int = 128, t_end: float = 0.6, amp: float = 130.0, *, riemann="hlld", recon="plm",
             predictor="full", scheme="cube", ncontrol=200, verbose=True, nstepmax=20000, precision="f64",
             params: Params | None = None):
    """glmmhd_turb.jl:2191-2241: driven turbulence with OU forcing and a CFL dt each step.
    The next step's dt comes from the CFL max folded into the previous step."""
    if params is None:
        p0_ = 1.0 * 1.0 ** 2 / 1.01
        p = Params(N=N, turb_rms=amp, switch_llf_dmin=0.01, switch_llf_pmin=0.01 * p0_)
    else:
        p = replace(params, N=N, turb_rms=amp)
    s = Solver(p, precision)
    u1, u2 = s.new_state(), s.new_state()
    p0 = p.rho0 * p.cs0 ** 2 / p.gamma
    b = tuple(p.b0 * c for c in p.b0_dir)
    s.ic_uniform(u1, p.rho0, p0, b, p.gamma)
    f = s.setup_forcing(p, amp)
    f.advance()
    t, step = 0.0, 0
    cmax = s.ctot_max(u1)
    while t < t_end and step < nstepmax:
        while t >= f.next_time:
            f.advance()
        dt = p.courant * dxof(p) / cmax
        dt = min(dt, f.next_time - t, t_end - t)
        s.step_plm(u1, u2, p, dt, t, do_turb=True, riemann=riemann, recon=recon, predictor=predictor,
                   scheme=scheme, want_cfl=True)
        u1, u2 = u2, u1
        t += dt
        step += 1
        cmax = s.cfl_max()
        if verbose and step % ncontrol == 0:
            d = s.diagnostics(u1)
            print(f"  step={step} t={t:.4f} dt={dt:.2e} Mach={d['mach']:.3f} divBmax={d['divbmax']:.2e} "
                  f"rhomin={d['rhomin']:.3e} finite={d['finite']}")
    d = s.diagnostics(u1)
    if verbose:
        print(f"TURB N={N} amp={amp:.0f}: steps={step} t={t:.3f} Mach={d['mach']:.3f} "
              f"divBmax={d['divbmax']:.2e} rhomin={d['rhomin']:.3e} finite={d['finite']}")
    return d, u1, s
