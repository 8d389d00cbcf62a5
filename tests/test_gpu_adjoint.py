"""GPU tests of the adjoint (glm_step_adjoint through the C ABI, and the checkpointed trajectory).

Pins (no reference counterpart exists):
* transpose identity <J v, w> = <v, J^T w> per step, with J v from the CPU oracle over dual
  numbers (liboracle_dual.so, an independent restatement of the step) and J^T w from the GPU
  reverse step, for all 32 shipped variants (plus two presets) and for the forcing amplitude;
* the whole checkpointed trajectory against central finite differences of the recorded map on
  random directions of the initial velocity and of the whole initial state (1e-6), and the
  forcing-amplitude derivative against the dual-number oracle over the same recorded steps;
* checkpointing: the gradient is bit-identical for any number of snapshots;
* the trajectory gradient at 32^3 against the dual-number oracle run over the same recorded steps.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import DualOracle, Oracle  # noqa: E402

from glmmhdturb_amd import Params, Solver, step_args  # noqa: E402
from glmmhdturb_amd.adjoint import (Trajectory, initial_state, solenoidal_velocity,  # noqa: E402
                                     velocity_gradient)
from test_gpu_parity import oracle_scheme, smooth_state  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


# Every reverse-step variant the library ships (2 predictors x 2 reconstructions x 2 limiters x 4
# solvers, with the cube preset's face guard and predictor revert), plus the multi preset and the
# spike preset's clamped faces without the revert.
SCHEMES = [dict(predictor=pr, recon=rc, slope=sl, riemann=rm, scheme="cube")
           for pr in ("full", "oned") for rc in ("plm", "ppm") for sl in ("moncen", "minmod")
           for rm in ("llf", "llf_bound", "hll", "hlld")] + [
    dict(predictor="full", recon="plm", slope="minmod", riemann="hll", scheme="multi"),
    dict(predictor="oned", recon="plm", slope="moncen", riemann="hlld", scheme="spike"),
]
SCHEME_IDS = ["-".join(v for v in k.values()) for k in SCHEMES]


@pytest.mark.parametrize("k", range(len(SCHEMES)), ids=SCHEME_IDS)
def test_step_transpose_identity_vs_dual_oracle(k):
    cfg = dict(SCHEMES[k])
    scheme = cfg.pop("scheme")
    orc, dorc = Oracle("f64"), DualOracle()
    nx, ny, nz = 12, 10, 8
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0, courant=0.4)
    s = Solver(p, "f64")
    u = smooth_state(orc, nx, ny, nz, seed=5 + k)
    dt = 0.4 * (1.0 / nx) / orc.ctot_max(u, p.gamma, p.smallr, p.smallr * p.smallc ** 2 / p.gamma)
    a = step_args(p, dt, 0.0, want_cfl=False, scheme=scheme, **cfg)
    if scheme == "spike":
        a.riemann = 3   # normal-only predictor and clamped faces of the spike preset, with HLLD
    sc = oracle_scheme(a)
    rng = np.random.default_rng(100 + k)
    v, w = rng.standard_normal(u.shape), rng.standard_normal(u.shape)
    _, jv = dorc.step(u, v, sc)
    lam0 = s.new_state()
    s.step_adjoint(dev(u), dev(w), lam0, a)
    jtw = lam0.cpu().numpy()
    lhs, rhs = float((jv * w).sum()), float((v * jtw).sum())
    print(f"[adjoint step] {SCHEMES[k]}: <Jv,w> {lhs:.15e} <v,J^T w> {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), np.abs(jv * w).sum() * 1e-3)
    s.close()


def test_step_forcing_amplitude_vs_dual_oracle():
    """d/d(scale) of one forced step with the field scaled: the dual oracle with the field as its
    own tangent against the GPU reverse step's forcing accumulator."""
    orc, dorc = Oracle("f64"), DualOracle()
    n = 12
    p = Params(N=n, gamma=5.0 / 3.0, courant=0.4, turb_gs=8, turb_kmin=1.0, turb_kmax=3.0, turb_T=0.5,
               comp_frac=0.0, turb_seed=3)
    s = Solver(p, "f64")
    s.setup_forcing(p, 40.0).advance()
    af = s.afield()
    u = smooth_state(orc, n, n, n, seed=9)
    dt = 0.3 * (1.0 / n) / orc.ctot_max(u, p.gamma, p.smallr, p.smallr * p.smallc ** 2 / p.gamma)
    a = step_args(p, dt, 0.2, do_turb=True, riemann="hlld", recon="ppm", want_cfl=False)
    sc = oracle_scheme(a, turb_gs=8)
    rng = np.random.default_rng(7)
    w = rng.standard_normal(u.shape)
    afn = af.cpu().numpy()
    _, jv = dorc.step(u, np.zeros_like(u), sc, afield=afn, dafield=afn)
    lam0 = s.new_state()
    s.adjoint_force_dot(reset=True)
    s.step_adjoint(dev(u), dev(w), lam0, a, af)
    fdot = s.adjoint_force_dot(reset=True)
    lhs = float((jv * w).sum())
    print(f"[adjoint forcing] <J_a a, w> {lhs:.15e} fdot {fdot:.15e}", flush=True)
    assert abs(lhs - fdot) <= 1e-10 * abs(lhs)
    s.close()


def make_traj(n=16, nsteps=12, window=6, amp=3e3, snapshots=3, seed=1, bpert=0.0, turb_T=0.5, ndt=100):
    p = Params(N=n, gamma=5.0 / 3.0, courant=0.3, b0=0.5, turb_T=turb_T, turb_Ndt=ndt, comp_frac=0.0,
               turb_gs=n, turb_kmin=1.0, turb_kmax=3.0, turb_seed=seed)
    tr = Trajectory(p, nsteps, window, amp, riemann="hlld", recon="ppm", snapshots=snapshots)
    v0 = dev(solenoidal_velocity(n, seed=seed + 10, kmax=4.0, rms=0.1))
    b = dev(solenoidal_velocity(n, seed=seed + 20, kmax=4.0, rms=bpert)) if bpert > 0 else None
    u0 = initial_state(tr.s, 1.0, 1.0, (0.5, 0.0, 0.0), p.gamma, v0, b)
    return tr, v0, u0, b


def fd_report(checks):
    for name, fd, ad in checks:
        print(f"[adjoint FD] {name}: FD {fd:.12e} adjoint {ad:.12e} rel {abs(fd - ad) / abs(fd):.2e}", flush=True)
    for name, fd, ad in checks:
        assert abs(fd - ad) <= 1e-6 * abs(fd), name


def test_trajectory_velocity_gradient_vs_central_differences():
    """Central differences of the recorded map along 3 velocity directions.  The initial field
    carries a solenoidal perturbation (rms 0.2) so that J = <E_mag> depends on the velocity at
    first order (with a uniform B0 the first-order change sum B0 . db vanishes and the derivative
    drowns in the rounding of J).  Density and pressure stay uniform along these directions."""
    tr, v0, u0, b = make_traj(bpert=0.2)
    J = tr.forward(u0)
    assert abs(tr.replay(u0) - J) <= 1e-14 * abs(J)
    lam0, damp = tr.pullback()
    g = velocity_gradient(lam0, 1.0, v0)
    eps = 1e-6
    checks = []
    for d in range(3):
        dv = dev(solenoidal_velocity(16, seed=40 + d, kmax=4.0, rms=0.1))
        jp = tr.replay(initial_state(tr.s, 1.0, 1.0, (0.5, 0.0, 0.0), tr.p.gamma, v0 + eps * dv, b))
        jm = tr.replay(initial_state(tr.s, 1.0, 1.0, (0.5, 0.0, 0.0), tr.p.gamma, v0 - eps * dv, b))
        checks.append((f"velocity {d}", (jp - jm) / (2 * eps), float((g * dv).sum().item())))
    fd_report(checks)
    tr.close()


def test_trajectory_gradient_32_random_directions():
    """The adjoint workload's check on 32 random velocity directions (solenoidal, shells 1-4,
    seeded; momentum and energy change as a velocity change at fixed density and pressure) from a
    developed 16^3 state (the end of a first run; a uniform density or pressure is a switch point
    of every limiter), over 12 recorded steps:
    * the adjoint equals the exact directional derivative of the recorded map, the dual-number
      oracle run over the same steps from the same state, to 1e-9 on every direction;
    * the central difference (eps = 1e-6) lies within the range of the exact derivatives at five
      points of its segment, start + t eps du for t = -1, -1/2, 0, 1/2, 1 (the dual oracle run from
      each point), widened by 1e-6 of the derivative plus the rounding of J over 2 eps.  The
      recorded map is piecewise smooth: limiter, PPM, wave-speed and HLLD decisions switch along the
      segment (the oracle's branch trace, the count of each outcome of every decision over the
      recorded steps, differs between the two ends on every one of these directions), so the
      central difference is the mean of the derivative along the segment and not the derivative at
      start.  Where the sampled derivatives agree to 1e-7, the central difference meets the
      derivative at start to 1e-6 itself."""
    pre, v0, u0, b = make_traj(bpert=0.2, nsteps=8, window=1)
    pre.forward(u0)
    start = pre.uN.clone()
    pre.close()
    tr, _, _, _ = make_traj(bpert=0.2)
    J = tr.forward(start)
    lam0, damp = tr.pullback()
    dorc, torc = DualOracle(), Oracle("trace")
    s0 = start.cpu().numpy()
    eps = 1e-6
    fields = [tr.fields[r.field].cpu().numpy() if r.field >= 0 else None for r in tr.records]
    schemes = [oracle_scheme(r.args, turb_gs=16) for r in tr.records]

    def dual_derivative(u, du):
        dual = 0.0
        for n in range(len(tr.records)):
            u, du = dorc.step(u, du, schemes[n], afield=fields[n])
            if tr.in_window(n + 1):
                dual += float((u[5:8] * du[5:8]).sum()) * tr.dV() / tr.window
        return dual

    def trace(u):
        torc.branch_trace(reset=True)
        for n in range(len(tr.records)):
            u = torc.step(u, schemes[n], afield=fields[n])
        return torc.branch_trace(reset=True)

    fd_noise = 4 * abs(J) * np.finfo(np.float64).eps / (2 * eps)   # a few rounding units of J
    tight = 0
    for d in range(32):
        dv = dev(solenoidal_velocity(16, seed=1000 + d, kmax=4.0, rms=0.1))
        du = torch.zeros_like(start)
        du[1:4] = start[0] * dv
        du[4] = (start[1:4] * dv).sum(dim=0)
        ad = float((lam0 * du).sum().item())
        du0 = du.cpu().numpy()
        dual = dual_derivative(s0.copy(), du0.copy())
        along = [dual] + [dual_derivative(s0 + t * eps * du0, du0.copy()) for t in (-1.0, -0.5, 0.5, 1.0)]
        fd = (tr.replay(start + eps * du) - tr.replay(start - eps * du)) / (2 * eps)
        sw = np.argwhere(trace(s0 + eps * du0) != trace(s0 - eps * du0))
        lo, hi = min(along), max(along)
        slack = 1e-6 * abs(dual) + fd_noise
        print(f"[adjoint 32 dirs] {d:2d}: dual {dual:.12e} adjoint {ad:.12e} (rel {abs(ad - dual) / abs(dual):.1e}) "
              f"FD {fd:.12e} (rel {abs(fd - dual) / abs(dual):.1e}); along the segment [{lo:.12e}, {hi:.12e}] "
              f"(spread {(hi - lo) / abs(dual):.1e}); switching (site, outcome) {sw.tolist()}", flush=True)
        assert abs(ad - dual) <= 1e-9 * abs(dual), d
        assert lo - slack <= fd <= hi + slack, d
        if hi - lo <= 1e-7 * abs(dual):
            assert abs(fd - dual) <= 1e-6 * abs(dual) + fd_noise, d
            tight += 1
    print(f"[adjoint 32 dirs] {tight} of 32 directions have derivatives along the segment within 1e-7", flush=True)
    tr.close()


def test_trajectory_state_and_amplitude_gradient():
    """Central differences along 2 smooth directions of the whole initial state (all nine
    variables), and the forcing-amplitude derivative, from a developed state: a uniform density or
    pressure is a switch point of every limiter (dl = dr = 0), so the trajectory starts from the
    end of a first run, where no variable is uniform.  Short forcing time (T = 0.05) so that the
    kick is ramped up within the run."""
    pre, v0, u0, b = make_traj(bpert=0.2, turb_T=0.05, ndt=10, amp=2e4, nsteps=8, window=1)
    pre.forward(u0)
    start = pre.uN.clone()
    pre.close()
    tr, _, _, _ = make_traj(turb_T=0.05, ndt=10, amp=2e4, nsteps=12, window=6)
    tr.forward(start)
    lam0, damp = tr.pullback()
    eps = 1e-6
    checks = []
    scale = torch.tensor([0.01, 0.01, 0.01, 0.01, 0.01, 0.01, 0.01, 0.01, 0.001], dtype=torch.float64).cuda()
    for d in range(2):
        du = torch.stack([dev(solenoidal_velocity(16, seed=60 + 3 * d + k % 3, kmax=3.0, rms=1.0)[k % 3])
                          for k in range(9)]) * scale[:, None, None, None]
        fd = (tr.replay(start + eps * du) - tr.replay(start - eps * du)) / (2 * eps)
        checks.append((f"state {d}", fd, float((lam0 * du).sum().item())))
    fd_report(checks)
    # The amplitude derivative is ~1e-6 of J here: a central difference either crosses limiter
    # switches (large step) or drowns in the rounding of J (small step).  It is checked against
    # the dual-number oracle run over the same recorded steps with the forcing fields as their
    # own tangents (d/d(scale) of the fields scaled by scale), which is exact.
    dorc = DualOracle()
    u = start.cpu().numpy()
    du = np.zeros_like(u)
    dJ = 0.0
    for n, r in enumerate(tr.records):
        af = tr.fields[r.field].cpu().numpy() if r.field >= 0 else None
        u, du = dorc.step(u, du, oracle_scheme(r.args, turb_gs=16), afield=af, dafield=af)
        if tr.in_window(n + 1):
            dJ += float((u[5:8] * du[5:8]).sum()) * tr.dV() / tr.window
    ad = damp * tr.amp
    print(f"[adjoint amplitude] dual {dJ:.15e} adjoint {ad:.15e}", flush=True)
    assert abs(dJ - ad) <= 1e-9 * abs(dJ)
    tr.close()


def test_checkpointing_is_bit_identical():
    """Any number of snapshots, and all states kept from the forward run, give the same bits."""
    grads = []
    for snaps in (0, 2, 40, "all"):
        tr, v0, u0, _ = make_traj(nsteps=9, window=4, snapshots=snaps)
        tr.forward(u0)
        lam0, damp = tr.pullback()
        grads.append((lam0.clone(), damp, tr.forward_steps, tr.reverse_steps))
        tr.close()
    for g in grads[1:]:
        assert torch.equal(g[0], grads[0][0]) and g[1] == grads[0][1]
    assert grads[0][3] == grads[1][3] == grads[2][3] == grads[3][3] == 9
    assert grads[0][2] > grads[1][2] >= grads[2][2] > grads[3][2] == 0


def test_trajectory_gradient_vs_dual_oracle_32():
    """dJ/dv0 . dv at 32^3 over 4 recorded steps: the GPU adjoint against the dual-number oracle
    run over the same recorded steps and forcing fields."""
    tr, v0, u0, _ = make_traj(n=32, nsteps=4, window=2, snapshots=1, seed=2)
    tr.forward(u0)
    lam0, _ = tr.pullback()
    dv = dev(solenoidal_velocity(32, seed=77, kmax=4.0, rms=0.1))
    ad = float((velocity_gradient(lam0, 1.0, v0) * dv).sum().item())
    dorc = DualOracle()
    u = u0.cpu().numpy()
    du = np.zeros_like(u)
    du[1:4] = dv.cpu().numpy()
    du[4] = (v0 * dv).sum(dim=0).cpu().numpy()
    dJ = 0.0
    for n, r in enumerate(tr.records):
        af = tr.fields[r.field].cpu().numpy() if r.field >= 0 else None
        u, du = dorc.step(u, du, oracle_scheme(r.args, turb_gs=32), afield=af)
        if tr.in_window(n + 1):
            dJ += float((u[5:8] * du[5:8]).sum()) * tr.dV() / tr.window
    print(f"[adjoint dual 32^3] dual {dJ:.15e} adjoint {ad:.15e}", flush=True)
    assert abs(dJ - ad) <= 1e-9 * abs(dJ)
    tr.close()


@pytest.mark.parametrize("k", range(len(SCHEMES)), ids=SCHEME_IDS)
def test_fused_back_half_matches_passes(k):
    """The reverse step's back half as one z-march kernel (glm_step_adjoint) against the same
    sums as separate passes (glm_step_adjoint_passes), on a grid with partial tiles: the sums run in
    the same order, so the results agree to the last bits the compiler's FMA choices leave."""
    cfg = dict(SCHEMES[k])
    scheme = cfg.pop("scheme")
    orc = Oracle("f64")
    nx, ny, nz = 20, 14, 11
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0, courant=0.4)
    s = Solver(p, "f64")
    u = smooth_state(orc, nx, ny, nz, seed=31 + k)
    dt = 0.4 * (1.0 / nx) / orc.ctot_max(u, p.gamma, p.smallr, p.smallr * p.smallc ** 2 / p.gamma)
    a = step_args(p, dt, 0.0, want_cfl=False, scheme=scheme, **cfg)
    if scheme == "spike":
        a.riemann = 3
    w = np.random.default_rng(200 + k).standard_normal(u.shape)
    out = []
    for passes in (False, True):
        lam0 = s.new_state()
        s.step_adjoint(dev(u), dev(w), lam0, a, passes=passes)
        out.append(lam0.cpu().numpy())
    d = np.abs(out[0] - out[1]).max()
    print(f"[adjoint back half] {SCHEMES[k]}: max |fused - passes| {d:.3e} (max |lam0| {np.abs(out[1]).max():.3e})",
          flush=True)
    assert d <= 1e-13 * np.abs(out[1]).max()
    s.close()


def test_reverse_step_zchunk_invariance():
    """The fused back half marches z-chunks that recompute their boundary planes: the reverse step
    gives the same bits for every chunking (chunks of 1, 3, 5 and 11 planes against the
    occupancy model's choice)."""
    orc = Oracle("f64")
    nx, ny, nz = 20, 14, 11
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0, courant=0.4)
    s = Solver(p, "f64")
    u = smooth_state(orc, nx, ny, nz, seed=77)
    w = np.random.default_rng(78).standard_normal(u.shape)
    dt = 0.4 * (1.0 / nx) / orc.ctot_max(u, p.gamma, p.smallr, p.smallr * p.smallc ** 2 / p.gamma)
    out = []
    for zc in (0, 1, 3, 5, 11):
        a = step_args(p, dt, 0.0, want_cfl=False, riemann="hlld", recon="ppm", zchunk=zc)
        lam0 = s.new_state()
        s.step_adjoint(dev(u), dev(w), lam0, a)
        out.append(lam0.cpu())
    for o in out[1:]:
        assert torch.equal(o, out[0])
    s.close()
