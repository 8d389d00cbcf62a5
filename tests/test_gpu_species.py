"""Passive species in the GLM march (glm_step_species; spike_mhd.cu -DNSP=2: the fractions ride the
mass flux) against the oracle's o_step_sp, and the 11-plane MarchBridge library.

Bars: FP64 relative L1 <= 1e-12 per plane (the species planes included) against the oracle; FP32
<= 2e-6 against the FP32 oracle; two complementary fractions keep summing to one (1e-12 in FP64)
over a forced 128^3 run; the MHD planes equal glm_step's without species."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from branch_states import KINDS, pocket_state  # noqa: E402
from test_gpu_parity import TOL64, oracle_scheme, rel_l1, smooth_state, to_dev, to_np  # noqa: E402
from test_oracle_species import fraction_field, with_species  # noqa: E402

from glmmhdturb_amd import Params, Solver, pfloor, step_args  # noqa: E402

G = 5.0 / 3.0
NX, NY, NZ = 24, 20, 16
COMBOS = [(p, r, m, "moncen") for p in ("full", "oned") for r in ("plm", "ppm", "ppm_spike") for m in ("hll", "hlld")]
COMBOS += [("full", "ppm", "llf", "minmod"), ("oned", "plm", "llf_bound", "minmod")]


def species_rel_l1(a, b, ch):
    err = list(rel_l1(a[:9], b[:9], ch=ch))
    for v in (9, 10):
        err.append(np.abs(a[v] - b[v]).sum() / np.abs(b[v]).sum())
    return np.array(err)


@pytest.mark.parametrize("pred,recon,riem,slope", COMBOS)
def test_species_step_vs_oracle_fp64(orc, pred, recon, riem, slope):
    p = Params(N=NX, nx=NX, ny=NY, nz=NZ, gamma=G, courant=0.4)
    s = Solver(p, "f64")
    u = with_species(orc, smooth_state(orc, NX, NY, NZ), fraction_field(NZ, NY, NX))
    dt = 0.3 * (1.0 / NX) / orc.ctot_max(u[:9], G, p.smallr, pfloor(p))
    a = step_args(p, dt, 0.0, riemann=riem, recon=recon, predictor=pred, slope=slope, want_cfl=False)
    ud, un = to_dev(s, u), s.new_state(11)
    uo = u.copy()
    for _ in range(3):
        s.step_species(ud, un, a, 2)
        ud, un = un, ud
        uo = orc.step_sp(uo, oracle_scheme(a), 2)
    err = species_rel_l1(to_np(ud), uo, a.ch)
    assert np.all(err <= TOL64), (pred, recon, riem, slope, err)
    s.close()


@pytest.mark.parametrize("kind,riem", [("void", "hlld"), ("floor", "hll"), ("degenerate", "hlld")])
def test_species_on_branch_states(orc, kind, riem):
    """Reverted predictors and faces and floored cells carry the species along (fraction kept)."""
    k = KINDS[kind]
    p = Params(N=NX, nx=NX, ny=NY, nz=NZ, gamma=G, courant=0.4, smallr=k["smallr"], smallc=k["smallc"],
               switch_llf_dmin=k["llf_dmin"], switch_llf_pmin=k["llf_pmin"])
    s = Solver(p, "f64")
    u = with_species(orc, pocket_state(orc, NX, NY, NZ, kind, gamma=G), fraction_field(NZ, NY, NX))
    dt = p.courant * (1.0 / NX) / orc.ctot_max(u[:9], G, p.smallr, pfloor(p))
    a = step_args(p, dt, 0.0, riemann=riem, recon="ppm", predictor="full", want_cfl=False)
    ud, un = to_dev(s, u), s.new_state(11)
    uo = u.copy()
    for _ in range(3):
        s.step_species(ud, un, a, 2)
        ud, un = un, ud
        uo = orc.step_sp(uo, oracle_scheme(a), 2)
    err = species_rel_l1(to_np(ud), uo, a.ch)
    assert np.all(err <= TOL64), (kind, err)
    s.close()


def test_species_step_vs_oracle_fp32(orc32, orc):
    p = Params(N=NX, nx=NX, ny=NY, nz=NZ, gamma=G, courant=0.4)
    s = Solver(p, "f32")
    u = with_species(orc, smooth_state(orc, NX, NY, NZ), fraction_field(NZ, NY, NX)).astype(np.float32)
    dt = 0.3 * (1.0 / NX) / orc.ctot_max(u[:9].astype(np.float64), G, p.smallr, pfloor(p))
    a = step_args(p, dt, 0.0, riemann="hlld", recon="ppm", predictor="full", want_cfl=False)
    ud, un = to_dev(s, u), s.new_state(11)
    uo = u.copy()
    for _ in range(3):
        s.step_species(ud, un, a, 2)
        ud, un = un, ud
        uo = orc32.step_sp(uo, oracle_scheme(a), 2)
    err = species_rel_l1(to_np(ud), uo, a.ch)
    assert np.all(err <= 2e-6), err
    s.close()


def test_species_do_not_change_the_mhd_planes(orc):
    p = Params(N=NX, nx=NX, ny=NY, nz=NZ, gamma=G, courant=0.4)
    s = Solver(p, "f64")
    u = with_species(orc, smooth_state(orc, NX, NY, NZ), fraction_field(NZ, NY, NX))
    a = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=False)
    ud, un = to_dev(s, u), s.new_state(11)
    s.step_species(ud, un, a, 2)
    vd, vn = to_dev(s, u[:9]), s.new_state()
    s.step(vd, vn, a)
    assert np.max(rel_l1(to_np(un)[:9], to_np(vn), ch=a.ch)) <= 1e-14
    s.close()


def test_species_sum_to_one_in_forced_turbulence():
    """128^3 FP64, OU-forced from the benchmark's state with a two-species split: after 20 steps
    the fractions still sum to one everywhere, and the split has moved."""
    N = 128
    p = Params(N=N, gamma=G, courant=0.3, turb_gs=64, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0, turb_T=0.5,
               turb_Ndt=100, turb_seed=1, b0=0.5)
    s = Solver(p, "f64")
    u, un = s.new_state(11), s.new_state(11)
    s.ic_gradient(u[:9], G)
    x = (torch.arange(N, device=u.device, dtype=torch.float64) + 0.5) / N
    x1 = (0.5 + 0.4 * torch.sin(2 * np.pi * x)).view(1, 1, N).expand(N, N, N)
    u[9].copy_(u[0] * x1)
    u[10].copy_(u[0] * (1.0 - x1))
    f = s.setup_forcing(p, 1.8e5)
    f.advance()
    t = 0.0
    for _ in range(20):
        while t >= f.next_time:
            f.advance()
        dt = min(0.3 / N / s.ctot_max(u[:9]), f.next_time - t)
        s.step_species(u, un, step_args(p, dt, t, do_turb=True, riemann="hlld", recon="ppm", want_cfl=False), 2)
        u, un = un, u
        t += dt
    tot = (u[9] + u[10]) / u[0]
    assert float((tot - 1.0).abs().max()) <= 1e-12
    assert float((u[9] / u[0] - x1).abs().max()) > 1e-4
    s.close()


def test_libmhd_nsp2_through_march_bridge(orc, orc32):
    """NSP=2 build_mhd.sh -> libmhd480_nsp2.so (spike_mhd.cu -DNSP=2): 11 planes through the
    MarchBridge mirror against the FP32 oracle's species step (spike scheme)."""
    import os
    from oracle import Scheme
    from glmmhdturb_amd import march_bridge as MB
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "glmmhdturb_amd")
    m = MB.open_lib(os.path.join(pkg, "libmhd480_nsp2.so"))
    assert m.NV == 11 and m.f_glm is not None and m.h.march_nsp() == 2
    N = 24
    MB.set_grid(m, N, N, N)
    MB.set_gamma(m, 5.0 / 3.0)
    dtdx, ch, fac = 0.04, 1.1, 0.98
    MB.set_dtdx(m, dtdx)
    MB.set_glm(m, ch, fac)
    u0 = with_species(orc, smooth_state(orc, N, N, N, seed=21), fraction_field(N, N, N)).astype(np.float32)
    q = [torch.from_numpy(u0[v].copy()).cuda() for v in range(11)]
    o = [torch.empty_like(x) for x in q]
    assert MB.run_(m, q, o, 3) >= 0.0
    g32 = float(np.float32(5.0 / 3.0))
    sc = Scheme(gamma=g32, smallr=1e-6, pfl=1e-6, dt=np.float32(dtdx), dx=1.0, ch=np.float32(ch),
                glm_fac=np.float32(fac), predictor="oned", recon="plm", slope="moncen", riemann="hll",
                mh_revert=False, face_guard="clamp", update_floor=False)
    uo = u0.copy()
    for _ in range(3):
        uo = orc32.step_sp(uo, sc, 2)
    got = np.stack([x.cpu().numpy() for x in q]).astype(np.float64)
    err = species_rel_l1(got, uo.astype(np.float64), ch)
    assert np.all(err <= 2e-6), err
