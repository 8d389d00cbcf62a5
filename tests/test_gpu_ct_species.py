"""Passive species in the fused CT march (glm_ct_step_species; spike_ctm.cu -DNSP=2: the fractions
ride the mass flux) against the oracle's o_ct_step_sp, and the 10-plane CT bridge library.

Bars: FP64 relative L1 <= 1e-12 per plane (the species planes included) against the oracle, with
the floor and the OU kick as well; FP32 <= 2e-6 against the FP32 oracle; the 8 CT planes equal
glm_ct_step's to 1e-13 (the rounding of a different instantiation); every z-chunking gives the
same bits; two complementary fractions keep summing to one (1e-12 in FP64) over forced steps at
64^3."""
import ctypes as C
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import CTScheme, OUForcing  # noqa: E402
from test_gpu_ct_march import ct_args, ct_state  # noqa: E402
from test_gpu_parity import TOL64, rel_l1, to_dev, to_np  # noqa: E402
from test_oracle_species import fraction_field  # noqa: E402

from glmmhdturb_amd import Params, Solver, _lib  # noqa: E402

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "glmmhdturb_amd")
G = 5.0 / 3.0


def with_species(u8, x1):
    rho = u8[0]
    return np.concatenate([u8, (rho * x1)[None], (rho * (1.0 - x1))[None]])


def ct_species_rel_l1(a, b):
    err = list(rel_l1(a[:8], b[:8], 8))
    for v in (8, 9):
        err.append(np.abs(a[v] - b[v]).sum() / np.abs(b[v]).sum())
    return np.array(err)


@pytest.mark.parametrize("riem,recon,fs,shape", [
    ("hlld", "plm", 0, (24, 20, 16)), ("hll", "plm", 1, (24, 20, 16)), ("llf", "ppm_spike", 0, (24, 20, 16)),
    ("hlld", "ppm_spike", 1, (20, 13, 11)), ("hll", "ppm_spike", 0, (17, 30, 9))])
def test_ct_species_vs_oracle_fp64(orc, riem, recon, fs, shape):
    nx, ny, nz = shape
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=G)
    s = Solver(p, "f64")
    u = with_species(ct_state(orc, nx, ny, nz), fraction_field(nz, ny, nx))
    dx = 1.0 / nx
    dt = 0.3 * dx / orc.ct_ctot_max(u[:8], G, 1e-6, 1e-6)
    a = ct_args(p, dt, riem, recon)
    a.update_floor = 0
    cs = CTScheme(gamma=G, smallr=1e-6, smallp=1e-6, dt=dt, dx=dx, face_slope=fs, riemann=riem, recon=recon)
    qd, qn = to_dev(s, u), s.new_state(10)
    q8, q8n = to_dev(s, u[:8].copy()), s.new_state(8)
    qo = u.copy()
    for _ in range(3):
        s.ct_step_species(qd, qn, a, 2, fs)
        qd, qn = qn, qd
        s.ct_step(q8, q8n, a, fs)
        q8, q8n = q8n, q8
        qo = orc.ct_step_sp(qo, cs, 2)
    err = ct_species_rel_l1(to_np(qd), qo)
    print(f"[ct species] {riem} {recon} fs={fs} {shape}: rel L1 {err.max():.2e}", flush=True)
    assert np.all(err <= TOL64), err
    # the species do not feed back: the CT planes equal glm_ct_step's up to the FMA contraction of a
    # kernel compiled with two more values per cell
    assert np.all(rel_l1(to_np(qd[:8]), to_np(q8), 8) <= 1e-13)
    s.close()


def test_ct_species_forced_floor_vs_oracle(orc):
    """The OU kick and the density floor in the fused update with species: a near-vacuum block is
    floored, and there rho X = smallr X_old."""
    N = 24
    p = Params(N=N, gamma=G, courant=0.3, turb_gs=16, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0,
               turb_T=0.5, turb_Ndt=100, turb_seed=5)
    s = Solver(p, "f64")
    f = s.setup_forcing(p, amp=50.0)
    fo = OUForcing(orc, g=16, kmin=1.0, kmax=3.0, T=0.5, ndt=100, comp_frac=0.0, amp=50.0, seed=5)
    u8 = ct_state(orc, N, N, N, seed=4)
    blk = (slice(8, 14), slice(6, 12), slice(4, 10))
    u8[0][blk] = 1e-9
    u8[1][blk] = 0.0; u8[2][blk] = 0.0; u8[3][blk] = 0.0
    u8[4][blk] = 1e-3 / (G - 1.0) + 0.5 * (u8[5][blk] ** 2 + u8[6][blk] ** 2 + u8[7][blk] ** 2)
    u = with_species(u8, fraction_field(N, N, N))
    dx = 1.0 / N
    qd, qn = to_dev(s, u), s.new_state(10)
    qo = u.copy()
    t = 0.0
    floored = 0
    for _ in range(4):
        while t >= f.next_time:
            f.advance(); fo.advance()
        dt = min(0.3 * dx / orc.ct_ctot_max(qo[:8], G, 1e-6, 1e-6), f.next_time - t)
        a = ct_args(p, dt, "hlld", "plm", do_turb=True, t=t + 0.05)
        a.update_floor = 1
        cs = CTScheme(gamma=G, smallr=a.smallr, smallp=1e-6, dt=dt, dx=dx, smallc=a.smallc,
                      turb_min_rho=a.turb_min_rho, ramp=a.ramp, riemann="hlld", recon="plm", do_turb=True,
                      turb_gs=16, update_floor=True)
        s.ct_step_species(qd, qn, a, 2, 0)
        qd, qn = qn, qd
        qo = orc.ct_step_sp(qo, cs, 2, afield=fo.afield)
        floored += int((qo[0] == a.smallr).sum())
        t += dt
    assert floored > 0
    err = ct_species_rel_l1(to_np(qd), qo)
    assert np.all(err <= TOL64), err
    s.close()


def test_ct_species_fp32_vs_oracle(orc, orc32):
    nx, ny, nz = 24, 20, 16
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=G)
    s = Solver(p, "f32")
    u = with_species(ct_state(orc, nx, ny, nz, seed=3), fraction_field(nz, ny, nx, seed=7)).astype(np.float32)
    dx = 1.0 / nx
    dt = float(np.float32(0.3 * dx / orc.ct_ctot_max(u[:8].astype(np.float64), G, 1e-6, 1e-6)))
    a = ct_args(p, dt, "hlld", "ppm_spike")
    a.update_floor = 0
    cs = CTScheme(gamma=float(np.float32(G)), smallr=1e-6, smallp=1e-6, dt=dt, dx=float(np.float32(dx)),
                  riemann="hlld", recon="ppm_spike")
    qd, qn = to_dev(s, u), s.new_state(10)
    qo = u.copy()
    for _ in range(3):
        s.ct_step_species(qd, qn, a, 2, 0)
        qd, qn = qn, qd
        qo = orc32.ct_step_sp(qo, cs, 2)
    err = ct_species_rel_l1(to_np(qd), qo.astype(np.float64))
    assert np.all(err <= 2e-6), err
    s.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ct_species_zchunk_invariance(orc, prec):
    nx, ny, nz = 24, 20, 30
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=G)
    s = Solver(p, prec)
    u = with_species(ct_state(orc, nx, ny, nz, seed=6), fraction_field(nz, ny, nx))
    dx = 1.0 / nx
    dt = 0.3 * dx / orc.ct_ctot_max(u[:8], G, 1e-6, 1e-6)
    qd = to_dev(s, u)
    outs = []
    for zc in (0, 1, 4, 7, 30):
        qn = s.new_state(10)
        s.ct_step_species(qd, qn, ct_args(p, dt, "hlld", "ppm_spike", zchunk=zc), 2, 0)
        outs.append(qn)
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    s.close()


def test_ct_species_sum_to_one_in_forced_turbulence(orc):
    """64^3 FP64, OU-forced, HLLD + PLM: two complementary fractions still sum to one everywhere
    after 30 steps, and the split has moved."""
    N = 64
    p = Params(N=N, gamma=G, courant=0.3, turb_gs=32, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0, turb_T=0.2,
               turb_Ndt=50, turb_seed=2)
    s = Solver(p, "f64")
    u8 = ct_state(orc, N, N, N, seed=12, b0=0.3)
    x1 = fraction_field(N, N, N, seed=9)
    q, qn = to_dev(s, with_species(u8, x1)), s.new_state(10)
    f = s.setup_forcing(p, 4e3)
    f.advance()
    t, dx = 0.0, 1.0 / N
    for _ in range(30):
        while t >= f.next_time:
            f.advance()
        cmax = s.ct_ctot_max(q[:8])
        dt = min(0.3 * dx / cmax, f.next_time - t)
        s.ct_step_species(q, qn, ct_args(p, dt, "hlld", "plm", do_turb=True, t=t), 2, 0)
        q, qn = qn, q
        t += dt
    tot = (q[8] + q[9]) / q[0]
    assert float((tot - 1.0).abs().max()) <= 1e-12
    assert float((q[8] / q[0] - torch.from_numpy(x1).cuda()).abs().max()) > 1e-4
    assert s.ct_divb_max(q[:8], dx) <= 1e-9
    s.close()


def test_ct_species_rejects_other_paths_and_counts(orc):
    """Species run on the fused march only; nsp must be 0 or 2."""
    n = 16
    p = Params(N=n, gamma=G)
    s = Solver(p, "f64")
    q, o = s.new_state(10), s.new_state(10)
    a = ct_args(p, 1e-3, "hlld", "plm")
    with pytest.raises(Exception):
        s.ct_step_species(q, o, a, 1, 0)
    qa, kq = s._planes(q, 10)
    oa, ko = s._planes(o, 10)
    assert s.lib.glm_ct_step_species(s.ctx, qa, oa, C.byref(a), 0, 3) != 0
    s.close()


def test_libct_nsp2_ct_run(orc, orc32):
    """NSP=2 build_ct.sh -> libct480_nsp2.so (spike_ctm.cu built with -DNSP=2): ct_run over 10
    planes against the FP32 oracle's species step (spike scheme: PLM MonCen, HLL, normal-only Hancock)."""
    lib = C.CDLL(os.path.join(PKG, "libct480_nsp2.so"))
    lib.march_nv.restype = C.c_int
    lib.march_nsp.restype = C.c_int
    lib.march_set_grid.restype = C.c_int
    lib.ct_run.restype = C.c_double
    lib.ct_run.argtypes = [C.POINTER(C.c_void_p)] * 3 + [C.c_int]
    assert lib.march_nv() == 10 and lib.march_nsp() == 2
    N = 24
    assert lib.march_set_grid(N, N, N) == 0
    g32 = float(np.float32(G))
    lib.march_set_gamma(C.c_float(g32))
    dtdx = float(np.float32(0.02))
    lib.march_set_dtdx(C.c_float(dtdx))
    u = with_species(ct_state(orc, N, N, N), fraction_field(N, N, N)).astype(np.float32)
    q = [torch.from_numpy(u[v].copy()).cuda() for v in range(10)]
    o = [torch.empty_like(x) for x in q]
    cc = [torch.empty_like(x) for x in q[:3]]
    qa, kq = _lib.plane_array([x.data_ptr() for x in q])
    oa, ko = _lib.plane_array([x.data_ptr() for x in o])
    ca, kc = _lib.plane_array([x.data_ptr() for x in cc])
    assert lib.ct_run(qa, oa, ca, 3) >= 0.0
    cs = CTScheme(gamma=g32, smallr=1e-6, smallp=1e-6, dt=dtdx, dx=1.0, riemann="hll", recon="plm", face_slope=0)
    uo = u.copy()
    for _ in range(3):
        uo = orc32.ct_step_sp(uo, cs, 2)
    got = np.stack([x.cpu().numpy() for x in q])
    assert np.all(ct_species_rel_l1(got, uo.astype(np.float64)) <= 2e-6)
