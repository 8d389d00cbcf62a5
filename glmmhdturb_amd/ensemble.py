"""Ensemble of GLM-MHD boxes on one GPU (libglmmhd ``glm_ens_*``).

Boxes of uneven size advance together.  A sweep is one step of every box -- each with its own dt
from its own CFL max and its own seeded OU forcing, i.e. ``run_turb``'s loop
(glmmhd_turb.jl:2191-2241) evaluated on the device -- and one march launch over all boxes'
tiles.  A sweep needs no host round trip, so it can be captured once into a HIP graph and
replayed.  The reference has no ensemble mode; each box here is the ``run_turb`` box of
``turb.Solver`` with the same parameters.
"""
from __future__ import annotations

import ctypes as C
import math
import random
from dataclasses import dataclass, field, replace

from . import _lib
from ._lib import GLMEnsBox, check
from .turb import NV, Params, Solver, dxof, step_args


@dataclass
class BoxSpec:
    """One box: grid, initial state (uniform, glmmhd_turb.jl:2164-2172) and forcing."""
    N: int
    nx: int | None = None
    ny: int | None = None
    nz: int | None = None
    boxlen: float = 1.0
    courant: float = 0.3
    rho0: float = 1.0
    p0: float = 1.0
    b0: tuple = (0.5, 0.0, 0.0)
    do_turb: bool = True
    amp: float = 1e5
    seed: int = 1
    turb_T: float = 0.5
    turb_Ndt: int = 100
    comp_frac: float = 0.0
    turb_gs: int = 64
    turb_kmin: float = 1.0
    turb_kmax: float = 3.0
    ch_scale: float = 0.25
    cp_coef: float = 0.18

    @property
    def shape(self):
        return (self.nx or self.N, self.ny or self.N, self.nz or self.N)

    def cells(self) -> int:
        nx, ny, nz = self.shape
        return nx * ny * nz


def box_params(spec: BoxSpec, scheme: Params) -> Params:
    """The single-box Params of a box (its turb.Solver and run_turb settings)."""
    nx, ny, nz = spec.shape
    return replace(scheme, N=nx, nx=nx, ny=ny, nz=nz, boxlen=spec.boxlen, courant=spec.courant,
                   rho0=spec.rho0, turb_T=spec.turb_T, turb_Ndt=spec.turb_Ndt, comp_frac=spec.comp_frac,
                   turb_seed=spec.seed, turb_gs=spec.turb_gs, turb_kmin=spec.turb_kmin,
                   turb_kmax=spec.turb_kmax, glm_ch_scale=spec.ch_scale, glm_cp_coef=spec.cp_coef,
                   turb_rms=spec.amp)


# Forcing amplitude that drives a box (rho = p = 1, gamma 5/3, |B0| = 0.5) to a saturated rms Mach
# ~1 with the 64^3 forcing grid, measured at 128^3 over t = 1..3: 1e5 -> 0.72, 1.6e5 -> 0.95,
# 2e5 -> 1.05, 2.6e5 -> 1.22, i.e. Mach ~ amp^0.6.  The field is amp/g^3 times a unit-variance
# spectrum sum, so amp scales with g^3 at a fixed acceleration.
AMP_MACH1_G64 = 1.8e5
MACH_AMP_EXPONENT = 0.6


def random_boxes(nbox: int = 64, seed: int = 2024, sizes=(32, 48, 64, 96, 128), mach=(0.3, 2.0),
                 beta=(0.5, 100.0), gamma: float = 5.0 / 3.0) -> list[BoxSpec]:
    """A seeded ensemble of cubic boxes: size uniform over `sizes`, target rms Mach uniform in
    `mach` (forcing amplitude ~ Mach^(1/0.6), the measured saturation law above), plasma
    beta uniform in `beta` (|B0| = sqrt(2 p0 / beta) along x), rho0 = p0 = 1, OU in shells 1-3
    with T = 0.5 on a min(64, N)^3 forcing grid, its own seed per box."""
    rng = random.Random(seed)
    out = []
    for b in range(nbox):
        n = rng.choice(sizes)
        m = rng.uniform(*mach)
        bt = rng.uniform(*beta)
        g = min(64, n)
        amp = AMP_MACH1_G64 * m ** (1.0 / MACH_AMP_EXPONENT) * (g / 64.0) ** 3
        out.append(BoxSpec(N=n, b0=(math.sqrt(2.0 / bt), 0.0, 0.0), amp=amp, seed=1000 + b, turb_gs=g))
    return out


class Ensemble:
    """glm_ens: boxes, their state buffers (two (9, nz, ny, nx) torch tensors each) and a
    turb.Solver per box for initial conditions and diagnostics."""

    def __init__(self, boxes: list[BoxSpec], scheme: Params | None = None, precision: str = "f64", *,
                 riemann="hlld", recon="ppm", predictor=None, slope="moncen", scheme_name="cube"):
        import torch
        self.lib = _lib.load()
        self.scheme = scheme or Params(gamma=5.0 / 3.0)
        self.boxes = list(boxes)
        self.prec = precision
        self.solvers, self.u0, self.u1 = [], [], []
        for spec in self.boxes:
            s = Solver(box_params(spec, self.scheme), precision)
            a, b = s.new_state(), s.new_state()
            s.ic_uniform(a, spec.rho0, spec.p0, spec.b0, self.scheme.gamma)
            self.solvers.append(s)
            self.u0.append(a)
            self.u1.append(b)
        torch.cuda.synchronize()
        arr = (GLMEnsBox * len(self.boxes))()
        for k, spec in enumerate(self.boxes):
            x = arr[k]
            x.nx, x.ny, x.nz = spec.shape
            x.do_turb = int(spec.do_turb)
            for v in range(NV):
                x.u0[v] = self.u0[k][v].data_ptr()
                x.u1[v] = self.u1[k][v].data_ptr()
            x.boxlen, x.courant, x.ch_scale, x.cp_coef = spec.boxlen, spec.courant, spec.ch_scale, spec.cp_coef
            x.g, x.ndt = spec.turb_gs, spec.turb_Ndt
            x.kmin, x.kmax, x.turb_T = spec.turb_kmin, spec.turb_kmax, spec.turb_T
            x.comp_frac, x.amp, x.seed = spec.comp_frac, spec.amp, spec.seed
        self.args = step_args(self.scheme, 1.0, 0.0, riemann=riemann, recon=recon, predictor=predictor,
                              slope=slope, scheme=scheme_name, do_turb=True)
        self.ens = self.lib.glm_ens_create(len(self.boxes), arr, C.byref(self.args), 8 if precision == "f64" else 4)
        if not self.ens:
            raise _lib.GLMError(self.lib.glm_error_string(self.lib.glm_last_error()).decode())
        self._arr = arr

    # ---- lifecycle ------------------------------------------------------------------------
    def close(self):
        if getattr(self, "ens", None):
            self.lib.glm_ens_destroy(self.ens)
            self.ens = None
        for s in getattr(self, "solvers", []):
            s.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nbox(self) -> int:
        return len(self.boxes)

    def cells(self) -> int:
        return sum(b.cells() for b in self.boxes)

    def stream(self):
        """The ensemble's HIP stream as a torch stream object."""
        import torch
        return torch.cuda.ExternalStream(self.lib.glm_ens_stream(self.ens))

    # ---- stepping -------------------------------------------------------------------------
    def reset(self):
        """Restart every box at t = 0 from the state now in its first buffer."""
        check(self.lib.glm_ens_reset(self.ens))

    def restart(self):
        """Re-apply each box's initial state to its first buffer and restart at t = 0."""
        import torch
        self.sync()
        for spec, s, u in zip(self.boxes, self.solvers, self.u0):
            s.ic_uniform(u, spec.rho0, spec.p0, spec.b0, self.scheme.gamma)
        torch.cuda.synchronize()
        self.reset()

    def step(self, nsweeps: int = 1):
        check(self.lib.glm_ens_step(self.ens, int(nsweeps)))

    def graph(self, on: bool = True):
        check(self.lib.glm_ens_graph(self.ens, int(on)))

    def sync(self):
        check(self.lib.glm_ens_sync(self.ens))

    @property
    def sweeps(self) -> int:
        return int(self.lib.glm_ens_sweeps(self.ens))

    def configure(self, zchunk: int = 0, persistent: int = 0, order: int = 1):
        check(self.lib.glm_ens_configure(self.ens, int(zchunk), int(persistent), int(order)))

    def tune(self, reps: int = 3):
        best = (C.c_int * 3)()
        ms = C.c_double()
        check(self.lib.glm_ens_tune(self.ens, int(reps), best, C.byref(ms)))
        return {"zchunk": best[0], "persistent": best[1], "order": best[2], "march_ms": ms.value}

    # ---- state and records ----------------------------------------------------------------
    def current(self, b: int):
        """The tensor holding box b's current state."""
        planes = (C.c_void_p * NV)()
        check(self.lib.glm_ens_current(self.ens, int(b), planes))
        return self.u0[b] if planes[0] == self.u0[b][0].data_ptr() else self.u1[b]

    def log(self, b: int, count: int) -> list[dict]:
        out = (C.c_double * (count * _lib.ENS_NLOG))()
        check(self.lib.glm_ens_log(self.ens, int(b), int(count), out))
        return [dict(zip(_lib.ENS_LOG_KEYS, out[k * _lib.ENS_NLOG:(k + 1) * _lib.ENS_NLOG])) for k in range(count)]

    def timing(self, count: int):
        """Per sweep (oldest first) and box: ms from the box's first tile start to its last tile end."""
        n = self.nbox
        out = (C.c_double * (count * n))()
        check(self.lib.glm_ens_timing(self.ens, int(count), out))
        return [list(out[k * n:(k + 1) * n]) for k in range(count)]

    def diagnostics(self, b: int) -> dict:
        return self.solvers[b].diagnostics(self.current(b))


def single_box_run(spec: BoxSpec, scheme: Params, nsteps: int, precision: str = "f64", *, riemann="hlld",
                   recon="ppm", predictor=None, slope="moncen", scheme_name="cube", log=None):
    """The same box alone through turb.Solver with run_turb's host loop (glmmhd_turb.jl:2213-2226).
    With `log` (the ensemble's step log of this box) the step's ch and psi decay are taken from it,
    so that the march of both runs sees identical scalars; dt is recomputed here."""
    import torch
    p = box_params(spec, scheme)
    s = Solver(p, precision)
    u1, u2 = s.new_state(), s.new_state()
    s.ic_uniform(u1, spec.rho0, spec.p0, spec.b0, scheme.gamma)
    f = s.setup_forcing(p, spec.amp) if spec.do_turb else None
    t, cmax = 0.0, s.ctot_max(u1)
    rec = []
    for k in range(nsteps):
        if f is not None:
            while t >= f.next_time:
                f.advance()
        dt = p.courant * dxof(p) / cmax
        if f is not None:
            dt = min(dt, f.next_time - t)
        kw = {}
        if log is not None:
            kw = dict(ch=log[k]["ch"], glm_fac=log[k]["glm_fac"])
        a = step_args(p, dt, t, do_turb=spec.do_turb, riemann=riemann, recon=recon, predictor=predictor,
                      slope=slope, scheme=scheme_name, want_cfl=True, **kw)
        rec.append({"t": t, "dt": dt, "ch": a.ch, "glm_fac": a.glm_fac, "ramp": a.ramp, "cmax": cmax})
        s.step(u1, u2, a)
        u1, u2 = u2, u1
        t += dt
        cmax = s.cfl_max()
    torch.cuda.synchronize()
    return s, u1, rec
