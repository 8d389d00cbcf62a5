"""Adjoint of a forced GLM-MHD trajectory (no reference counterpart: the reference has no
derivatives).  The counterpart of a ChainRulesCore ``rrule`` / ``torch.autograd.Function``:

* ``Trajectory.forward(u0)`` runs ``run_turb``'s loop (glmmhd_turb.jl:2213-2226) for ``nsteps``
  steps from u0, records every step's scalars (dt, ch, psi decay, ramp) and every forcing field
  it used, and returns the objective J = time average over the last ``window`` steps of the
  magnetic energy sum 0.5 |B|^2 dV.
* ``Trajectory.pullback(jbar)`` returns (jbar dJ/dU0, jbar dJ/damp): the reverse sweep runs one
  ``glm_step_adjoint`` per step over states recomputed from checkpoints with binomial
  checkpointing (``snapshots`` stored states), or, with ``snapshots="all"``, over the states the
  forward run kept in device memory (one state per step: 151 MB at 128^3 FP64, so 200 steps fit
  many times over in an MI355X's 288 GB) with nothing recomputed.

The derivative is that of the recorded map: dt and the forcing realisation are held at their
recorded values (the CFL dt is a max over cells, not a smooth function of the state).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from functools import lru_cache

from .turb import NV, Params, Solver, dxof, step_args


@lru_cache(maxsize=None)
def _schedule(n: int, s: int) -> tuple[int, int]:
    """(forward steps, first split) of the cheapest recursive checkpointing schedule that reverses
    n steps from a stored start state with s further checkpoints: store the state after m steps,
    reverse the right n - m steps with s - 1 checkpoints, then the left m steps with s.  With no
    checkpoint left, each step's state is recomputed from the start.  This is the binomial
    schedule of Griewank and Walther (Algorithm 799, 'revolve'), found here by exhaustive
    minimisation over m."""
    if n <= 1:
        return 0, 0
    if s <= 0:
        return n * (n - 1) // 2, 0
    best, arg = None, 1
    for m in range(1, n):
        c = m + _schedule(n - m, s - 1)[0] + _schedule(m, s)[0]
        if best is None or c < best:
            best, arg = c, m
    return best, arg


def split_point(n: int, s: int) -> int:
    return _schedule(n, s)[1]


def recompute_steps(n: int, s: int) -> int:
    """Forward steps the reverse sweep recomputes for n steps with s checkpoints."""
    return _schedule(n, s)[0]


@dataclass
class StepRecord:
    args: object        # GLMParams of the step
    field: int          # index of the forcing field the step used (-1: none)


class Trajectory:
    def __init__(self, p: Params, nsteps: int, window: int, amp: float, *, riemann="hlld", recon="ppm",
                 predictor=None, slope="moncen", scheme="cube", snapshots=16, do_turb: bool = True):
        import torch
        self.torch = torch
        self.p, self.nsteps, self.window, self.amp = p, int(nsteps), int(window), float(amp)
        self.kw = dict(riemann=riemann, recon=recon, predictor=predictor, slope=slope, scheme=scheme)
        self.store_all = snapshots == "all"
        self.snapshots = 0 if self.store_all else int(snapshots)
        self.states = []
        self.do_turb = do_turb
        self.s = Solver(p, "f64")
        if do_turb:
            self.s.setup_forcing(p, amp)
        self.records: list[StepRecord] = []
        self.fields = []
        self.u0 = None
        self.forward_steps = 0
        self.reverse_steps = 0

    # ---- objective ------------------------------------------------------------------------
    def dV(self) -> float:
        return dxof(self.p) ** 3

    def in_window(self, n: int) -> bool:
        """State U_n (after n steps) enters J."""
        return self.nsteps - self.window < n <= self.nsteps

    def emag(self, u) -> float:
        return float(0.5 * (u[5:8].double() ** 2).sum().item()) * self.dV()

    # ---- forward ----------------------------------------------------------------------------
    def forward(self, u0):
        """Run and record the trajectory from u0 (not modified); returns J."""
        torch = self.torch
        s, p = self.s, self.p
        self.u0 = u0.clone()
        self.records, self.fields = [], []
        self.states = []
        u, w = u0.clone(), s.new_state()
        f = s.forcing
        if f is not None:
            s.setup_forcing(p, self.amp)
            f = s.forcing
        t, cmax = 0.0, s.ctot_max(u)
        J = 0.0
        for n in range(self.nsteps):
            fi = -1
            if f is not None:
                while t >= f.next_time:
                    f.advance()
                    self.fields.append(s.afield())
                fi = len(self.fields) - 1
            dt = p.courant * dxof(p) / cmax
            if f is not None:
                dt = min(dt, f.next_time - t)
            a = step_args(p, dt, t, do_turb=f is not None, want_cfl=True, **self.kw)
            self.records.append(StepRecord(a, fi))
            if self.store_all and n > 0:
                self.states.append(u.clone())      # U_n (U_0 is self.u0)
            s.step(u, w, a)
            u, w = w, u
            t += dt
            cmax = s.cfl_max()
            if self.in_window(n + 1):
                J += self.emag(u) / self.window
        torch.cuda.synchronize()
        self.uN = u.clone()
        self.J = J
        return J

    def replay(self, u0, amp_scale: float = 1.0):
        """J of the recorded map (recorded dt, ch, psi decay, ramp) from u0 with the recorded
        forcing fields scaled by amp_scale: the function the adjoint differentiates."""
        s = self.s
        u, w = u0.clone(), s.new_state()
        J = 0.0
        for n, r in enumerate(self.records):
            if r.field >= 0:
                s.set_afield(self.fields[r.field] * amp_scale)
            s.step(u, w, r.args)
            u, w = w, u
            if self.in_window(n + 1):
                J += self.emag(u) / self.window
        self.torch.cuda.synchronize()
        return J

    # ---- reverse ----------------------------------------------------------------------------
    def _advance(self, u, n0: int, n1: int):
        """State after step n1 - 1 from state u after n0 steps (u is not modified)."""
        s = self.s
        a, b = u.clone(), s.new_state()
        for n in range(n0, n1):
            r = self.records[n]
            if r.field >= 0:
                s.set_afield(self.fields[r.field])
            s.step(a, b, r.args)
            a, b = b, a
            self.forward_steps += 1
        return a

    def _reverse_one(self, n: int, un, lam):
        """Adjoint of U_n from the adjoint of U_{n+1} (lam), plus dJ/dU_n."""
        r = self.records[n]
        out = self.s.new_state()
        self.s.step_adjoint(un, lam, out, r.args, self.fields[r.field] if r.field >= 0 else None)
        self.reverse_steps += 1
        if self.in_window(n):
            out[5:8] += un[5:8] * (self.dV() / self.window)
        return out

    def _reverse(self, a: int, b: int, ua, s: int, lam):
        if b - a == 1:
            return self._reverse_one(a, ua, lam)
        if s <= 0:
            for i in range(b - 1, a - 1, -1):
                lam = self._reverse_one(i, self._advance(ua, a, i) if i > a else ua, lam)
            return lam
        m = a + split_point(b - a, s)
        um = self._advance(ua, a, m)
        lam = self._reverse(m, b, um, s - 1, lam)
        del um
        return self._reverse(a, m, ua, s, lam)

    def pullback(self, jbar: float = 1.0):
        """(jbar dJ/dU0 as a (9, nz, ny, nx) tensor, jbar dJ/damp)."""
        if self.u0 is None:
            raise RuntimeError("call forward() first")
        s = self.s
        self.forward_steps = self.reverse_steps = 0
        lam = s.new_state()
        if self.in_window(self.nsteps):
            lam[5:8] = self.uN[5:8] * (self.dV() / self.window)
        s.adjoint_force_dot(reset=True)
        if self.store_all:
            for n in range(self.nsteps - 1, -1, -1):
                lam = self._reverse_one(n, self.states[n - 1] if n > 0 else self.u0, lam)
            lam0 = lam
        else:
            lam0 = self._reverse(0, self.nsteps, self.u0, self.snapshots, lam)
        fdot = s.adjoint_force_dot(reset=True)
        self.torch.cuda.synchronize()
        damp = fdot / self.amp if self.amp != 0 else 0.0
        return lam0 * jbar, damp * jbar

    def close(self):
        self.states = []
        self.s.close()


def solenoidal_velocity(N: int, seed: int, kmin: float = 1.0, kmax: float = 4.0, rms: float = 0.1):
    """A seeded solenoidal Gaussian velocity field on an N^3 grid with power in the shells
    kmin <= |k| <= kmax, normalised to the given rms: (3, N, N, N) numpy array (z, y, x)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(N, 1.0 / N)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    shell = (kk >= kmin) & (kk <= kmax)
    v = np.empty((3, N, N, N))
    comps = [rng.standard_normal((N, N, N)) for _ in range(3)]
    hat = [np.fft.fftn(c) * shell for c in comps]
    with np.errstate(invalid="ignore", divide="ignore"):
        kdot = (KX * hat[0] + KY * hat[1] + KZ * hat[2]) / np.where(kk > 0, kk ** 2, 1.0)
    proj = [hat[0] - KX * kdot, hat[1] - KY * kdot, hat[2] - KZ * kdot]
    for c in range(3):
        v[c] = np.real(np.fft.ifftn(proj[c]))
    v *= rms / np.sqrt((v ** 2).sum(axis=0).mean())
    return v


def initial_state(solver: Solver, rho0: float, p0: float, b0, gamma: float, v, b=None):
    """Uniform state (glmmhd_turb.jl:2164-2172) with velocity field v and, optionally, a magnetic
    perturbation b added to B0 ((3, nz, ny, nx) tensors)."""
    u = solver.new_state()
    solver.ic_uniform(u, rho0, p0, b0, gamma)
    u[1:4] = rho0 * v
    u[4] += 0.5 * rho0 * (v ** 2).sum(dim=0)
    if b is not None:
        bb = u[5:8].clone()
        u[5:8] = bb + b
        u[4] += 0.5 * ((bb + b) ** 2 - bb ** 2).sum(dim=0)
    return u


def velocity_gradient(lam0, rho0: float, v):
    """dJ/dv of J(U0(v)) for initial_state: rho0 (lam_m + lam_E v)."""
    return rho0 * (lam0[1:4] + lam0[4] * v)
