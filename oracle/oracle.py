"""ctypes wrapper over the C oracle (oracle/_build/liboracle_{f64,f32,dual}.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, where it is the reference the GPU results are compared with (or, in the
cpu_baseline leg, the CPU path that is timed).  The product package never imports it.

Every function mirrors one in glm_oracle.c / ou_oracle.c / ct_oracle.c; see those files for the
reference lines each one restates.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
NV = 9

RIEMANN = {"llf": 0, "llf_bound": 1, "hll": 2, "hlld": 3}
RECON = {"plm": 0, "ppm": 1, "ppm_spike": 2}
PREDICTOR = {"full": 0, "oned": 1}
SLOPE = {"moncen": 0, "minmod": 1}
FACE_GUARD = {"none": 0, "iflo": 1, "clamp": 2}


def build() -> None:
    """Compile the oracle (gcc, both precisions)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OScheme(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "smallr", "smallc", "pfl", "dt", "dx", "ch", "glm_fac", "llf_dmin", "llf_pmin",
        "turb_min_rho", "ramp", "boxlen")] + [(n, C.c_int) for n in (
        "predictor", "recon", "slope", "riemann", "mh_revert", "face_guard", "update_floor", "do_turb",
        "turb_gs", "pad_")]


class OCTScheme(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "smallr", "smallp", "dt", "dx", "smallc", "turb_min_rho", "ramp")] + [(n, C.c_int) for n in (
        "face_slope", "riemann", "recon", "do_turb", "turb_gs", "update_floor")]


class OHyScheme(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "dtdx", "smallr", "smallp")] + [
        ("fallback", C.c_int), ("fb_dmin", C.c_double), ("fb_pmin", C.c_double), ("nsp", C.c_int), ("cma", C.c_int),
        ("logsp", C.c_int), ("ppm", C.c_int), ("donor", C.c_int)]


@dataclass
class HydroScheme:
    """The 5-variable hydro march (spike_25d.cu:180 constants; LLF fallback :335-376), with nsp = 2
    passive scalars (-DSCALARS) and their flux riding the mass flux with cma (-DCMA)."""
    gamma: float = 1.4
    dtdx: float = 0.05
    smallr: float = 1e-6
    smallp: float = 1e-6
    fallback: bool = False
    fb_dmin: float = 1e-3
    fb_pmin: float = 1e-3
    nsp: int = 0
    cma: bool = False
    logsp: bool = False        # the scalar planes hold log2 X (-DU16SP)
    ppm: bool = False          # parabolic edges (-DPPM1D), HLL without the fallback
    donor: bool = False        # donor-cell 32 x 8 tile seams (-DDONOR), HLL without the fallback

    def to_c(self) -> OHyScheme:
        return OHyScheme(self.gamma, self.dtdx, self.smallr, self.smallp, int(self.fallback), self.fb_dmin,
                         self.fb_pmin, self.nsp, int(self.cma), int(self.logsp), int(self.ppm), int(self.donor))


@dataclass
class Scheme:
    """Scheme switches + run scalars, with names from Params (glmmhd_turb.jl:28-52)."""
    gamma: float = 5.0 / 3.0
    smallr: float = 1e-6
    smallc: float = 1e-4
    pfl: float | None = None          # default pfloor = smallr*smallc^2/gamma (glmmhd_turb.jl:55)
    dt: float = 1e-3
    dx: float = 1.0 / 64
    ch: float = 1.0
    glm_fac: float = 1.0
    llf_dmin: float = 0.0
    llf_pmin: float = 0.0
    turb_min_rho: float = 1e-5
    ramp: float = 0.0
    boxlen: float = 1.0
    predictor: str = "full"
    recon: str = "plm"
    slope: str = "moncen"
    riemann: str = "hlld"
    mh_revert: bool = True
    face_guard: str = "iflo"
    update_floor: bool = True
    do_turb: bool = False
    turb_gs: int = 64

    def to_c(self) -> OScheme:
        pfl = self.pfl if self.pfl is not None else self.smallr * self.smallc ** 2 / self.gamma
        return OScheme(self.gamma, self.smallr, self.smallc, pfl, self.dt, self.dx, self.ch, self.glm_fac,
                       self.llf_dmin, self.llf_pmin, self.turb_min_rho, self.ramp, self.boxlen,
                       PREDICTOR[self.predictor], RECON[self.recon], SLOPE[self.slope], RIEMANN[self.riemann],
                       int(self.mh_revert), FACE_GUARD[self.face_guard], int(self.update_floor),
                       int(self.do_turb), self.turb_gs, 0)


@dataclass
class CTScheme:
    gamma: float = 5.0 / 3.0
    smallr: float = 1e-6
    smallp: float = 1e-6
    dt: float = 1e-3
    dx: float = 1.0 / 64
    smallc: float = 1e-4
    turb_min_rho: float = 1e-5
    ramp: float = 0.0
    face_slope: int = 0
    riemann: str = "hll"
    recon: str = "plm"
    do_turb: bool = False
    turb_gs: int = 64
    update_floor: bool = False

    def to_c(self) -> OCTScheme:
        return OCTScheme(self.gamma, self.smallr, self.smallp, self.dt, self.dx, self.smallc, self.turb_min_rho,
                         self.ramp, self.face_slope, RIEMANN[self.riemann], RECON[self.recon], int(self.do_turb),
                         self.turb_gs, int(self.update_floor))


_LIBS: dict = {}


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One precision of the oracle: dtype float64 ('f64') or float32 ('f32')."""

    def __init__(self, prec: str = "f64"):
        """prec: 'f64', 'f32', or 'trace' (FP64 with the branch trace of o_trace_get)."""
        path = os.path.join(BUILD, f"liboracle_{prec}.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.dtype = np.float32 if prec == "f32" else np.float64
        self.real = C.c_float if prec == "f32" else C.c_double
        L, rl, vp, i = self.lib, self.real, C.c_void_p, C.c_int
        sig = {
            "o_cons2prim": (None, [vp, vp, rl, rl, rl]),
            "o_prim2cons": (None, [vp, vp, rl]),
            "o_fast_speed": (rl, [vp, rl, rl]),
            "o_phys_flux_x": (None, [vp, vp, rl]),
            "o_rot_to": (None, [vp, vp, i]),
            "o_rot_flux_from": (None, [vp, vp, i]),
            "o_moncen": (rl, [rl, rl]),
            "o_minmod": (rl, [rl, rl]),
            "o_hancock": (None, [vp, vp, vp, vp, rl, rl, vp]),
            "o_hanc1d": (None, [vp, vp, i, rl, rl, rl, rl, vp]),
            "o_ppm_mono": (None, [rl, rl, rl, vp, vp]),
            "o_ppm_edges": (None, [rl, rl, rl, vp, vp]),
            "o_spj": (i, [rl, rl, rl]),
            "o_ppm1_spike": (rl, [rl, rl, rl, rl]),
            "o_glm_pair": (None, [rl, rl, rl, rl, rl, vp, vp]),
            "o_hlld_x": (None, [vp, vp, rl, rl, rl, rl, vp]),
            "o_riemann": (None, [vp, vp, i, i, rl, rl, rl, rl, rl, rl, vp]),
            "o_turb_interp": (None, [vp, i, i, i, i, rl, rl, rl, rl, rl, vp]),
            "o_update_cell": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp, i, i, i, i, i, i]),
            "o_recon_cell_faces": (None, [vp, i, i, i, i, i, i, vp, vp, vp]),
            "o_branch_hits": (C.c_longlong, [i]),
            "o_branch_hits_reset": (None, []),
            "o_trace_get": (None, [vp]),
            "o_trace_reset": (None, []),
            "o_step": (i, [vp, vp, i, i, i, vp, vp, vp]),
            "o_step_scratch_reals": (C.c_size_t, [i, i, i]),
            "o_step_sp": (i, [vp, vp, i, i, i, i, vp, vp, vp]),
            "o_step_sp_scratch_reals": (C.c_size_t, [i, i, i, i]),
            "o_ctot_max": (C.c_double, [vp, i, i, i, C.c_double, C.c_double, C.c_double]),
            "o_fast_max": (C.c_double, [vp, i, i, i, C.c_double, C.c_double, C.c_double]),
            "o_diagnostics": (None, [vp, i, i, i, C.c_double, C.c_double, vp]),
            "o_ic_ot": (None, [vp, i, i, i, C.c_double]),
            "o_ic_uniform": (None, [vp, i, i, i] + [C.c_double] * 6),
            "o_ic_gradient": (None, [vp, i, i, i, C.c_double]),
            "o_philox4x32_10": (None, [vp, vp, vp]),
            "o_ou_gauss6": (None, [C.c_uint64, C.c_uint64, C.c_uint32, vp]),
            "o_ou_init": (None, [i, C.c_double, C.c_double, vp, vp]),
            "o_ou_step": (None, [i, vp, vp, vp, vp] + [C.c_double] * 4 + [C.c_uint64, C.c_uint64]),
            "o_ou_afield": (None, [i, vp, vp, C.c_double, vp, vp, vp, vp]),
            "o_ct_step": (i, [vp, vp, i, i, i, vp, vp, vp]),
            "o_ct_step_scratch_reals": (C.c_size_t, [i, i, i]),
            "o_ct_step_sp": (i, [vp, vp, i, i, i, i, vp, vp, vp]),
            "o_ct_step_sp_scratch_reals": (C.c_size_t, [i, i, i, i]),
            "o_ct_divb_max": (C.c_double, [vp, i, i, i, C.c_double]),
            "o_ct_ctot_max": (C.c_double, [vp, i, i, i, C.c_double, C.c_double, C.c_double]),
            "o_ct_face_flux": (None, [vp, vp, vp, vp, rl, i, vp, vp]),
            "o_ct_dflux": (None, [vp, i, rl, vp]),
            "o_ct_cellprim": (None, [vp, vp, rl, rl, rl, vp]),
            "oh_c2p": (None, [vp, vp, vp]),
            "oh_moncen": (rl, [rl, rl, rl]),
            "oh_edge": (None, [vp, vp, rl, vp, vp]),
            "oh_flux": (None, [vp, i, vp, vp]),
            "oh_tocons": (None, [vp, vp, vp]),
            "oh_hanc1d": (None, [vp, vp, i, vp, vp]),
            "oh_hll": (None, [vp, vp, i, vp, vp]),
            "oh_riemann_fb": (None, [vp, vp, i, vp, vp]),
            "oh_step": (i, [vp, vp, i, i, i, vp, vp]),
            "oh_step_prim": (i, [vp, vp, i, i, i, vp, vp]),
            "oh_step_scratch_reals": (C.c_size_t, [i, i, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args

    # ---- helpers ------------------------------------------------------------
    def arr(self, x, n=None):
        a = np.ascontiguousarray(np.asarray(x, dtype=self.dtype))
        if n is not None:
            assert a.size == n, (a.size, n)
        return a

    def vec(self, n=NV):
        return np.zeros(n, dtype=self.dtype)

    # ---- point physics --------------------------------------------------------
    def cons2prim(self, c, gamma, smallr, pfl):
        c = self.arr(c, NV); q = self.vec()
        self.lib.o_cons2prim(_ptr(c), _ptr(q), gamma, smallr, pfl)
        return q

    def prim2cons(self, q, gamma):
        q = self.arr(q, NV); c = self.vec()
        self.lib.o_prim2cons(_ptr(q), _ptr(c), gamma)
        return c

    def fast_speed(self, q, gamma, bn):
        return self.lib.o_fast_speed(_ptr(self.arr(q, NV)), gamma, bn)

    def phys_flux_x(self, q, gamma):
        q = self.arr(q, NV); f = self.vec()
        self.lib.o_phys_flux_x(_ptr(q), _ptr(f), gamma)
        return f

    def rot_to(self, q, d):
        q = self.arr(q, NV); r = self.vec()
        self.lib.o_rot_to(_ptr(q), _ptr(r), d)
        return r

    def rot_flux_from(self, f, d):
        f = self.arr(f, NV); r = self.vec()
        self.lib.o_rot_flux_from(_ptr(f), _ptr(r), d)
        return r

    def moncen(self, a, b):
        return self.lib.o_moncen(a, b)

    def minmod(self, a, b):
        return self.lib.o_minmod(a, b)

    def hancock(self, m0, sx, sy, sz, dtdx, gamma):
        out = self.vec()
        args = [self.arr(v, NV) for v in (m0, sx, sy, sz)]
        self.lib.o_hancock(*[_ptr(a) for a in args], dtdx, gamma, _ptr(out))
        return out

    def hanc1d(self, c0, s, d, dtdx, gamma, smallr, pfl):
        c0 = self.arr(c0, NV); s = self.arr(s, NV); out = self.vec()
        self.lib.o_hanc1d(_ptr(c0), _ptr(s), d, dtdx, gamma, smallr, pfl, _ptr(out))
        return out

    def ppm_mono(self, ql, q0, qr):
        lo = self.real(); hi = self.real()
        self.lib.o_ppm_mono(ql, q0, qr, C.addressof(lo), C.addressof(hi))
        return lo.value, hi.value

    def ppm_edges(self, qm, q0, qp):
        lo = self.real(); hi = self.real()
        self.lib.o_ppm_edges(qm, q0, qp, C.addressof(lo), C.addressof(hi))
        return lo.value, hi.value

    def spj(self, a, b, c):
        return bool(self.lib.o_spj(a, b, c))

    def ppm1_spike(self, qm, q0, qp, sgn):
        return self.lib.o_ppm1_spike(qm, q0, qp, sgn)

    def glm_pair(self, bnL, bnR, psiL, psiR, ch):
        a = self.real(); b = self.real()
        self.lib.o_glm_pair(bnL, bnR, psiL, psiR, ch, C.addressof(a), C.addressof(b))
        return a.value, b.value

    def hlld_x(self, L, R_, gamma, bn, fbn=0.0, fpsi=0.0):
        L = self.arr(L, NV); R_ = self.arr(R_, NV); f = self.vec()
        self.lib.o_hlld_x(_ptr(L), _ptr(R_), gamma, bn, fbn, fpsi, _ptr(f))
        return f

    def riemann(self, L, R_, d, kind, gamma, ch, smallr, pfl, llf_dmin=0.0, llf_pmin=0.0):
        L = self.arr(L, NV); R_ = self.arr(R_, NV); f = self.vec()
        k = RIEMANN[kind] if isinstance(kind, str) else kind
        self.lib.o_riemann(_ptr(L), _ptr(R_), d, k, gamma, ch, smallr, pfl, llf_dmin, llf_pmin, _ptr(f))
        return f

    def update_cell(self, r, fxl, fxh, fyl, fyh, fzl, fzh, sc: "Scheme", afield=None, ijk=(0, 0, 0), nxyz=(1, 1, 1)):
        """o_update_cell: the fused update epilogue of one cell (returns the new state)."""
        r = self.arr(r, NV).copy()
        fl = [self.arr(f, NV) for f in (fxl, fxh, fyl, fyh, fzl, fzh)]
        af = None if afield is None else self.arr(afield)
        c = sc.to_c()
        self.lib.o_update_cell(_ptr(r), *[_ptr(f) for f in fl], C.byref(c), None if af is None else _ptr(af),
                               *ijk, *nxyz)
        return r

    def recon_faces(self, prim, i, j, k, sc: "Scheme"):
        """o_recon_cell_faces: (fm, fp), each (3, 9), the face states of cell (i, j, k)."""
        prim = self.arr(prim)
        nv, nz, ny, nx = prim.shape
        fm = self.vec(3 * NV); fp = self.vec(3 * NV)
        c = sc.to_c()
        self.lib.o_recon_cell_faces(_ptr(prim), nx, ny, nz, i, j, k, C.byref(c), _ptr(fm), _ptr(fp))
        return fm.reshape(3, NV), fp.reshape(3, NV)

    BRANCHES = ("hlld_degenerate", "mh_revert", "iflo_revert", "update_floor", "llf_switch")

    def branch_hits(self, reset=False):
        """How often each guarded branch fired since the last reset (name -> count)."""
        h = {b: int(self.lib.o_branch_hits(w)) for w, b in enumerate(self.BRANCHES)}
        if reset:
            self.lib.o_branch_hits_reset()
        return h

    TRACE_SITES = 32

    def branch_trace(self, reset=False):
        """(32, 4) counts of each outcome of each decision site (liboracle_trace.so only)."""
        out = np.zeros(self.TRACE_SITES * 4, dtype=np.int64)
        self.lib.o_trace_get(_ptr(out))
        if reset:
            self.lib.o_trace_reset()
        return out.reshape(self.TRACE_SITES, 4)

    def turb_interp(self, afield, gs, i, j, k, dx, lx, ly, lz, ramp):
        a = self.arr(afield, 3 * gs ** 3); out = self.vec(3)
        self.lib.o_turb_interp(_ptr(a), gs, i, j, k, dx, lx, ly, lz, ramp, _ptr(out))
        return out

    # ---- whole step ------------------------------------------------------------
    def new_state(self, nx, ny, nz, nv=NV):
        return np.zeros((nv, nz, ny, nx), dtype=self.dtype)

    def step(self, u, sc: Scheme, afield=None, out=None):
        nv, nz, ny, nx = u.shape
        u = self.arr(u)
        out = self.new_state(nx, ny, nz) if out is None else out
        scratch = np.empty(self.lib.o_step_scratch_reals(nx, ny, nz), dtype=self.dtype)
        cs = sc.to_c()
        af = None if afield is None else self.arr(afield)
        self.lib.o_step(_ptr(u), _ptr(out), nx, ny, nz, C.byref(cs), None if af is None else _ptr(af), _ptr(scratch))
        return out

    def step_sp(self, u, sc: Scheme, nsp: int, afield=None):
        """One step of a (9 + nsp)-plane state: the GLM-MHD planes and nsp passive species rho X
        (o_step_sp)."""
        nv, nz, ny, nx = u.shape
        assert nv == 9 + nsp
        u = self.arr(u)
        out = self.new_state(nx, ny, nz, nv)
        scratch = np.empty(self.lib.o_step_sp_scratch_reals(nx, ny, nz, nsp), dtype=self.dtype)
        cs = sc.to_c()
        af = None if afield is None else self.arr(afield)
        rc = self.lib.o_step_sp(_ptr(u), _ptr(out), nx, ny, nz, nsp, C.byref(cs), None if af is None else _ptr(af),
                                _ptr(scratch))
        if rc:
            raise ValueError(f"o_step_sp: nsp must be 0..2, not {nsp}")
        return out

    def ctot_max(self, u, gamma, smallr, pfl):
        nv, nz, ny, nx = u.shape
        return self.lib.o_ctot_max(_ptr(self.arr(u)), nx, ny, nz, gamma, smallr, pfl)

    def fast_max(self, u, gamma, smallr, pfl):
        nv, nz, ny, nx = u.shape
        return self.lib.o_fast_max(_ptr(self.arr(u)), nx, ny, nz, gamma, smallr, pfl)

    def diagnostics(self, u, dx, smallr=1e-6):
        nv, nz, ny, nx = u.shape
        out = np.zeros(10, dtype=np.float64)
        self.lib.o_diagnostics(_ptr(self.arr(u)), nx, ny, nz, dx, smallr, _ptr(out))
        keys = ("mass", "ke2", "divbmax", "rhomin", "finite", "ekin", "emag", "cross_helicity", "etot", "psi2")
        return dict(zip(keys, out.tolist()))

    def ic_ot(self, nx, ny, nz, gamma):
        u = self.new_state(nx, ny, nz)
        self.lib.o_ic_ot(_ptr(u), nx, ny, nz, gamma)
        return u

    def ic_uniform(self, nx, ny, nz, rho0, p0, b0, gamma):
        u = self.new_state(nx, ny, nz)
        self.lib.o_ic_uniform(_ptr(u), nx, ny, nz, rho0, p0, b0[0], b0[1], b0[2], gamma)
        return u

    def ic_gradient(self, nx, ny, nz, gamma):
        u = self.new_state(nx, ny, nz)
        self.lib.o_ic_gradient(_ptr(u), nx, ny, nz, gamma)
        return u

    # ---- OU forcing --------------------------------------------------------------
    def philox(self, ctr, key):
        c = np.asarray(ctr, dtype=np.uint32); k = np.asarray(key, dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        self.lib.o_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
        return out

    def ou_gauss6(self, seed, step, mode):
        g = np.zeros(6, dtype=np.float64)
        self.lib.o_ou_gauss6(seed, step, mode, _ptr(g))
        return g

    def ou_init(self, g, kmin, kmax):
        sqp = np.zeros(g ** 3); unitk = np.zeros(3 * g ** 3)
        self.lib.o_ou_init(g, kmin, kmax, _ptr(sqp), _ptr(unitk))
        return sqp, unitk

    def ou_step(self, g, fre, fim, sqp, unitk, alpha, beta, comp, sol, seed, step):
        self.lib.o_ou_step(g, _ptr(fre), _ptr(fim), _ptr(sqp), _ptr(unitk), alpha, beta, comp, sol, seed, step)

    def ou_afield(self, g, fre, fim, amp):
        cs, sn = twiddles(g)
        work = np.zeros(4 * g ** 3)
        af = np.zeros(3 * g ** 3, dtype=self.dtype)
        self.lib.o_ou_afield(g, _ptr(fre), _ptr(fim), amp, _ptr(cs), _ptr(sn), _ptr(work), _ptr(af))
        return af

    # ---- CT -----------------------------------------------------------------------
    def ct_step(self, q, sc: CTScheme, afield=None):
        nv, nz, ny, nx = q.shape
        q = self.arr(q)
        out = np.zeros_like(q)
        scratch = np.empty(self.lib.o_ct_step_scratch_reals(nx, ny, nz), dtype=self.dtype)
        cs = sc.to_c()
        af = None if afield is None else self.arr(afield)
        self.lib.o_ct_step(_ptr(q), _ptr(out), nx, ny, nz, C.byref(cs), None if af is None else _ptr(af), _ptr(scratch))
        return out

    def ct_step_sp(self, q, sc: CTScheme, nsp: int, afield=None):
        """One CT step of an (8 + nsp)-plane state: the CT planes and nsp passive species rho X
        (o_ct_step_sp)."""
        nv, nz, ny, nx = q.shape
        assert nv == 8 + nsp
        q = self.arr(q)
        out = np.zeros_like(q)
        scratch = np.empty(self.lib.o_ct_step_sp_scratch_reals(nx, ny, nz, nsp), dtype=self.dtype)
        cs = sc.to_c()
        af = None if afield is None else self.arr(afield)
        rc = self.lib.o_ct_step_sp(_ptr(q), _ptr(out), nx, ny, nz, nsp, C.byref(cs), None if af is None else _ptr(af),
                                   _ptr(scratch))
        if rc:
            raise ValueError(f"o_ct_step_sp: nsp must be 0..2, not {nsp}")
        return out

    def ct_divb_max(self, q, dx):
        nv, nz, ny, nx = q.shape
        return self.lib.o_ct_divb_max(_ptr(self.arr(q)), nx, ny, nz, dx)

    def ct_ctot_max(self, q, gamma, smallr, smallp):
        nv, nz, ny, nx = q.shape
        return self.lib.o_ct_ctot_max(_ptr(self.arr(q)), nx, ny, nz, gamma, smallr, smallp)

    # ---- 5-variable hydro march ----------------------------------------------------
    def _hy(self, name, *vecs, extra=(), sc: HydroScheme):
        cs = sc.to_c()
        nv = 5 + sc.nsp
        out = self.vec(7)
        args = [_ptr(np.pad(self.arr(v, nv), (0, 7 - nv))) for v in vecs]
        getattr(self.lib, name)(*args, *extra, _ptr(out), C.byref(cs))
        return out[:nv]

    def hydro_c2p(self, c, sc: HydroScheme):
        return self._hy("oh_c2p", c, sc=sc)

    def hydro_tocons(self, q, sc: HydroScheme):
        return self._hy("oh_tocons", q, sc=sc)

    def hydro_flux(self, q, d, sc: HydroScheme):
        return self._hy("oh_flux", q, extra=(d,), sc=sc)

    def hydro_hanc1d(self, m, s, d, sc: HydroScheme):
        return self._hy("oh_hanc1d", m, s, extra=(d,), sc=sc)

    def hydro_hll(self, L, R_, d, sc: HydroScheme):
        return self._hy("oh_hll", L, R_, extra=(d,), sc=sc)

    def hydro_riemann_fb(self, L, R_, d, sc: HydroScheme):
        return self._hy("oh_riemann_fb", L, R_, extra=(d,), sc=sc)

    def hydro_edge(self, b, s, sign, sc: HydroScheme):
        cs = sc.to_c()
        nv = 5 + sc.nsp
        out = self.vec(7)
        b7, s7 = np.pad(self.arr(b, nv), (0, 7 - nv)), np.pad(self.arr(s, nv), (0, 7 - nv))
        self.lib.oh_edge(_ptr(b7), _ptr(s7), sign, _ptr(out), C.byref(cs))
        return out[:nv]

    def hydro_moncen(self, a, b, c):
        return self.lib.oh_moncen(a, b, c)

    def hydro_step(self, u, sc: HydroScheme):
        """One step of the march on a (5 + nsp, nz, ny, nx) conserved state."""
        nv, nz, ny, nx = u.shape
        assert nv == 5 + sc.nsp, nv
        u = self.arr(u)
        out = np.zeros_like(u)
        scratch = np.empty(self.lib.oh_step_scratch_reals(nx, ny, nz), dtype=self.dtype)
        cs = sc.to_c()
        self.lib.oh_step(_ptr(u), _ptr(out), nx, ny, nz, C.byref(cs), _ptr(scratch))
        return out

    def hydro_step_prim(self, q, sc: HydroScheme):
        """The same step from primitive planes [rho, vx, vy, vz, p] (rho, p floored as the tile's
        make_prim does); returns the conserved result (oh_step_prim)."""
        nv, nz, ny, nx = q.shape
        assert nv == 5 + sc.nsp, nv
        q = self.arr(q)
        out = np.zeros_like(q)
        scratch = np.empty(self.lib.oh_step_scratch_reals(nx, ny, nz), dtype=self.dtype)
        cs = sc.to_c()
        self.lib.oh_step_prim(_ptr(q), _ptr(out), nx, ny, nz, C.byref(cs), _ptr(scratch))
        return out


# ---- compact storage codecs of the hydro march (spike_25d.cu:157-168), float32 arithmetic ----
_F = np.float32


def dec_log2(u):
    """uint16 -> log2 X over [-110, 0]"""
    return _F(-110) + np.asarray(u).astype(np.float32) * (_F(110) / _F(65535))


def enc_log2(l2):
    t = (np.asarray(l2, dtype=np.float32) + _F(110)) * (_F(65535) / _F(110))
    t = np.minimum(np.maximum(t, _F(0)), _F(65535))
    return np.floor(t + _F(0.5)).astype(np.uint16)


def dec_vel(u):
    """uint16 -> velocity over [-8, 8]"""
    return _F(-8) + np.asarray(u).astype(np.float32) * (_F(16) / _F(65535))


def enc_vel(x):
    t = (np.minimum(np.maximum(np.asarray(x, dtype=np.float32), _F(-8)), _F(8)) + _F(8)) * (_F(65535) / _F(16))
    return np.floor(t + _F(0.5)).astype(np.uint16)


def dec_log2s(u):
    """uint16 -> log2 of rho or p over [-32, 32]"""
    return _F(-32) + np.asarray(u).astype(np.float32) * (_F(64) / _F(65535))


def enc_log2s(l2):
    t = (np.asarray(l2, dtype=np.float32) + _F(32)) * (_F(65535) / _F(64))
    t = np.minimum(np.maximum(t, _F(0)), _F(65535))
    return np.floor(t + _F(0.5)).astype(np.uint16)


def cons_to_stored_prim(U, smallr, smallp, gamma):
    """The store of the primitive formats (spike_25d.cu:499-511): rho as is, v = m / max(rho,
    smallr), p = max((gamma - 1)(E - rho v^2 / 2), smallp), in float32."""
    U = np.asarray(U, dtype=np.float32)
    ir = _F(1) / np.maximum(U[0], _F(smallr))
    v = U[1:4] * ir
    ke = _F(0.5) * U[0] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    p = np.maximum(_F(gamma - 1.0) * (U[4] - ke), _F(smallp))
    return np.stack([U[0], v[0], v[1], v[2], p])


def twiddles(g: int):
    m = np.arange(g, dtype=np.float64)
    return np.cos(2.0 * math.pi * m / g), np.sin(2.0 * math.pi * m / g)


class DualOracle:
    """The FP64 GLM step over forward-mode dual numbers (liboracle_dual.so): one step maps
    (U, dU) to (L(U), J_L(U) dU) with the step's scalars (dt, ch, psi decay, ramp) fixed.  The
    forcing field may carry a tangent (e.g. afield / amp for the derivative in the amplitude)."""

    def __init__(self):
        path = os.path.join(BUILD, "liboracle_dual.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        vp, i = C.c_void_p, C.c_int
        self.lib.od_step.argtypes = [vp, vp, i, i, i, vp, vp, vp]
        self.lib.od_step.restype = i
        self.lib.od_step_scratch_pairs.argtypes = [i, i, i]
        self.lib.od_step_scratch_pairs.restype = C.c_size_t

    def step(self, u, du, sc: "Scheme", afield=None, dafield=None):
        nv, nz, ny, nx = u.shape
        x = np.ascontiguousarray(np.stack([np.asarray(u, np.float64), np.asarray(du, np.float64)], axis=-1))
        out = np.empty_like(x)
        af = None
        if afield is not None:
            da = np.zeros_like(afield, dtype=np.float64) if dafield is None else dafield
            af = np.ascontiguousarray(np.stack([np.asarray(afield, np.float64), np.asarray(da, np.float64)], axis=-1))
        scratch = np.empty(2 * self.lib.od_step_scratch_pairs(nx, ny, nz), dtype=np.float64)
        cs = sc.to_c()
        self.lib.od_step(_ptr(x), _ptr(out), nx, ny, nz, C.cast(C.pointer(cs), C.c_void_p),
                         None if af is None else _ptr(af), _ptr(scratch))
        return out[..., 0].copy(), out[..., 1].copy()


class OUForcing:
    """Oracle OU forcing state (glmmhd_turb.jl:2117-2161 Forcing / ou_advance!)."""

    def __init__(self, orc: Oracle, g=64, kmin=1.0, kmax=None, T=0.3, ndt=100, comp_frac=0.5, amp=1.0, seed=1):
        self.o, self.g = orc, g
        self.kmin, self.kmax = kmin, (g // 2 if kmax is None else kmax)
        self.T, self.ndt, self.comp, self.amp, self.seed = T, ndt, comp_frac, amp, seed
        self.sqp, self.unitk = orc.ou_init(g, self.kmin, self.kmax)
        self.fre = np.zeros(3 * g ** 3); self.fim = np.zeros(3 * g ** 3)
        self.nstep = 0
        self.next_time = 0.0
        self.afield = np.zeros(3 * g ** 3, dtype=orc.dtype)

    def advance(self):
        dt = self.T / self.ndt
        alpha = 1.0 - dt / self.T
        beta = math.sqrt(2.0 * dt / self.T)
        self.o.ou_step(self.g, self.fre, self.fim, self.sqp, self.unitk, alpha, beta, self.comp, 1.0 - self.comp,
                       self.seed, self.nstep)
        self.nstep += 1
        self.afield = self.o.ou_afield(self.g, self.fre, self.fim, self.amp)
        self.next_time += dt
