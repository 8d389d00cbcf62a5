"""The hydro march on the compact storage formats (glm_hydro_step_compact: spike_25d.cu -DF16STATE,
-DU16STATE, -DU16SP, :64-95) against the oracle: the stored codes are decoded in numpy (the codecs of
:157-168 restated in oracle.py), the FP32 oracle steps the decoded state (oh_step_prim for the
primitive formats, log2 species for U16SP), and its result is encoded again.

Bars: the GPU computes in FP32 with its own contraction, so a result that lands within rounding of
a code boundary may take the neighbouring code: every code within one step (f16: one ulp, uint16:
one code) of the oracle's, at least 99% of them identical; the FP32 hydro planes of U16SP within
2e-6 relative L1.  Each step starts from the GPU's previous output, decoded, so the comparison does
not drift."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import HydroScheme, dec_log2, dec_log2s, dec_vel, enc_log2, enc_log2s, enc_vel, cons_to_stored_prim  # noqa: E402
from test_hydro_compact_oracle import hydro_state  # noqa: E402
from test_gpu_hydro import voids  # noqa: E402

from glmmhdturb_amd import Params, Solver, hydro_args  # noqa: E402

G = 1.4
SR = SP = 1e-6


def solver(nx, ny, nz):
    return Solver(Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=G), "f32")


def planes(arrs, dtype):
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda").view(dtype) if a.dtype.itemsize == 2
            else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda") for a in arrs]


def host(ts):
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in ts]


def prim_of(u):
    rho = u[0]
    v = u[1:4] / rho
    p = (G - 1.0) * (u[4] - 0.5 * rho * (v ** 2).sum(axis=0))
    return np.stack([rho, v[0], v[1], v[2], p])


def code_agreement(got, want, f16=False):
    """Largest code distance and the share of identical codes."""
    if f16:
        g = got.view(np.uint16).astype(np.int64)
        w = want.view(np.uint16).astype(np.int64)
        # f16 codes of one sign are ordered as integers; a sign flip only happens through 0
        g = np.where(g & 0x8000, -(g & 0x7FFF), g)
        w = np.where(w & 0x8000, -(w & 0x7FFF), w)
    else:
        g, w = got.astype(np.int64), want.astype(np.int64)
    return int(np.abs(g - w).max()), float((g == w).mean())


@pytest.mark.parametrize("fallback", [False, True])
def test_f16_primitive_storage(orc32, fallback):
    nx, ny, nz = 40, 20, 12
    s = solver(nx, ny, nz)
    q = prim_of(hydro_state(40)[:, :nz, :ny, :nx]).astype(np.float16)
    args = hydro_args(0.05, G, fallback=(1.05, 1e-3) if fallback else None)
    sc = HydroScheme(gamma=G, dtdx=0.05, fallback=fallback, fb_dmin=1.05, fb_pmin=1e-3)
    qd = planes(list(q), torch.float16)
    od = [torch.empty_like(x) for x in qd]
    worst, share = 0, 1.0
    for _ in range(3):
        cur = np.stack(host(qd))
        s.hydro_step_compact(qd, od, args, "f16_prim")
        got = np.stack(host(od))
        want = cons_to_stored_prim(orc32.hydro_step_prim(cur.astype(np.float32), sc), SR, SP, G).astype(np.float16)
        d, sh = code_agreement(got, want, f16=True)
        worst, share = max(worst, d), min(share, sh)
        qd, od = od, qd
    print(f"[f16 prim] fallback={fallback}: largest code distance {worst}, identical {share:.5f}", flush=True)
    assert worst <= 1 and share >= 0.99
    s.close()


def u16_prim_encode(prim):
    return np.stack([enc_log2s(np.log2(np.maximum(prim[0], np.float32(SR)))), enc_vel(prim[1]), enc_vel(prim[2]),
                     enc_vel(prim[3]), enc_log2s(np.log2(prim[4]))])


def u16_prim_decode(codes):
    return np.stack([np.exp2(dec_log2s(codes[0])), dec_vel(codes[1]), dec_vel(codes[2]), dec_vel(codes[3]),
                     np.exp2(dec_log2s(codes[4]))]).astype(np.float32)


def test_u16_primitive_storage(orc32):
    nx, ny, nz = 36, 24, 10
    s = solver(nx, ny, nz)
    q = u16_prim_encode(prim_of(hydro_state(36)[:, :nz, :ny, :nx]).astype(np.float32))
    args = hydro_args(0.05, G)
    sc = HydroScheme(gamma=G, dtdx=0.05)
    qd = planes(list(q), torch.int16)
    od = [torch.empty_like(x) for x in qd]
    worst, share = 0, 1.0
    for _ in range(3):
        cur = np.stack(host(qd)).view(np.uint16)
        s.hydro_step_compact(qd, od, args, "u16_prim")
        got = np.stack(host(od)).view(np.uint16)
        U = orc32.hydro_step_prim(u16_prim_decode(cur), sc)
        want = u16_prim_encode(cons_to_stored_prim(U, SR, SP, G))
        d, sh = code_agreement(got, want)
        worst, share = max(worst, d), min(share, sh)
        qd, od = od, qd
    print(f"[u16 prim]: largest code distance {worst}, identical {share:.5f}", flush=True)
    assert worst <= 1 and share >= 0.99
    s.close()


@pytest.mark.parametrize("fallback", [False, True])
def test_u16_log2_species(orc32, fallback):
    nx, ny, nz = 40, 16, 12
    s = solver(nx, ny, nz)
    u = (voids(nx, ny, nz, 5) if fallback else hydro_state(40)[:, :nz, :ny, :nx]).astype(np.float32)
    z, y, x = np.meshgrid(*(np.arange(n) / n for n in (nz, ny, nx)), indexing="ij")
    x1 = 0.3 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * z)
    x2 = 1e-20 * (1.0 + 0.5 * np.cos(2 * np.pi * y))          # a trace species, 20 dex down
    codes = [enc_log2(np.log2(x1.astype(np.float32))), enc_log2(np.log2(x2.astype(np.float32)))]
    args = hydro_args(0.05, G, fallback=(1e-3, 1e-3) if fallback else None)
    sc = HydroScheme(gamma=G, dtdx=0.05, nsp=2, cma=True, logsp=True, fallback=fallback)
    qd = planes(list(u) + codes, torch.int16)
    od = [torch.empty_like(t) for t in qd]
    worst, share, hydro_err = 0, 1.0, 0.0
    for _ in range(3):
        cur = host(qd)
        s.hydro_step_compact(qd, od, args, "u16_log2_species")
        got = host(od)
        qin = np.stack(cur[:5] + [dec_log2(cur[5].view(np.uint16)), dec_log2(cur[6].view(np.uint16))])
        U = orc32.hydro_step(qin.astype(np.float32), sc)
        rinv = np.float32(1) / np.maximum(U[0], np.float32(SR))
        want = [enc_log2(np.log2(np.maximum(U[v] * rinv, np.float32(7.7e-34)))) for v in (5, 6)]
        for v in (5, 6):
            d, sh = code_agreement(got[v].view(np.uint16), want[v - 5])
            worst, share = max(worst, d), min(share, sh)
        hydro_err = max(hydro_err, max(np.abs(got[v] - U[v]).sum() / np.abs(U[v]).sum() for v in range(5)))
        qd, od = od, qd
    print(f"[u16 log2 species] fallback={fallback}: hydro rel L1 {hydro_err:.2e}, largest code distance {worst}, "
          f"identical {share:.5f}", flush=True)
    assert hydro_err <= 2e-6 and worst <= 1 and share >= 0.99
    s.close()


def test_compact_formats_reject_bad_arguments():
    s = solver(16, 16, 16)
    q = [torch.zeros((16, 16, 16), dtype=torch.float16, device="cuda") for _ in range(5)]
    o = [torch.zeros_like(x) for x in q]
    with pytest.raises(ValueError):
        s.hydro_step_compact(q, o, hydro_args(0.05, G), "f8")
    with pytest.raises(ValueError):
        s.hydro_step_compact(q[:4], o[:4], hydro_args(0.05, G), "f16_prim")
    with pytest.raises(ValueError):        # 16-bit planes where U16SP takes float32 hydro planes
        s.hydro_step_compact(q + q[:2], o + o[:2], hydro_args(0.05, G), "u16_log2_species")
    s.close()
    s64 = Solver(Params(N=16, gamma=G), "f64")
    from glmmhdturb_amd._lib import plane_array
    import ctypes as C
    qa, kq = plane_array([x.data_ptr() for x in q])
    oa, ko = plane_array([x.data_ptr() for x in o])
    assert s64.lib.glm_hydro_step_compact(s64.ctx, qa, oa, C.byref(hydro_args(0.05, G)), 1) != 0   # FP32 only
    s64.close()
