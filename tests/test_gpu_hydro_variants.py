"""The parabolic-edge (recon PPM_SPIKE: spike_25d.cu -DPPM1D, :467-483) and donor-cell
(glm_hydro_step_donor: -DDONOR, :438-454) builds of the hydro march against the CPU oracle
(HydroScheme.ppm / .donor in oracle/hydro_oracle.c).

Bars: FP64 relative L1 <= 1e-12 per variable after several steps, FP32 <= 2e-6; chunkings of z
bit-identical.  The donor seams are fixed by the global cell index (x faces at i % 32 == 0, y faces
at j % 8 == 0, and the periodic faces), so grids with partial tiles and grids smaller than one tile
are cases of their own."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import HydroScheme  # noqa: E402
from glmmhdturb_amd import hydro_args  # noqa: E402
from test_gpu_hydro import G, rel_l1, run_oracle, solver  # noqa: E402
from test_hydro_oracle import smooth_hydro  # noqa: E402
from test_hydro_variants_oracle import periodic_state  # noqa: E402


def with_species(u, seed):
    rng = np.random.default_rng(seed)
    x0 = 0.5 + 0.3 * np.tanh(rng.normal(0, 1, u[0].shape))
    return np.concatenate([u, (u[0] * x0)[None], (u[0] * (1 - x0))[None]])


def run(s, u, args, steps, nsp=0, cma=False, donor=False):
    a, b = s.new_state(5 + nsp), s.new_state(5 + nsp)
    a.copy_(torch.from_numpy(np.ascontiguousarray(u)).to(a.device, a.dtype))
    for _ in range(steps):
        s.hydro_step(a, b, args, nsp=nsp, cma=cma, donor=donor)
        a, b = b, a
    torch.cuda.synchronize()
    return a.cpu().numpy().astype(np.float64)


def errs(got, ref):
    return np.array([np.abs(got[v] - ref[v]).sum() / max(np.abs(ref[v]).sum(), 1e-300) for v in range(len(ref))])


GRIDS = [(32, 8, 6), (70, 20, 6), (40, 20, 12), (7, 5, 3)]   # full tiles, partial tiles, smaller than a tile


@pytest.mark.parametrize("variant", ["ppm", "donor"])
@pytest.mark.parametrize("grid", GRIDS)
def test_variant_fp64_vs_oracle(orc, variant, grid):
    nx, ny, nz = grid
    u = periodic_state(nx, ny, nz, seed=nx + ny)
    sc = HydroScheme(gamma=G, dtdx=0.1, ppm=variant == "ppm", donor=variant == "donor")
    s = solver(nx, ny, nz)
    print(f"[hydro {variant} f64] grid {grid}", flush=True)
    got = run(s, u, hydro_args(0.1, G, ppm=variant == "ppm"), 5, donor=variant == "donor")
    ref = run_oracle(orc, u, sc, 5)
    err = errs(got, ref)
    print(f"[hydro {variant} f64] grid {grid}: rel-L1 {err.max():.2e}", flush=True)
    assert np.all(err <= 1e-12), err
    # and the variant is not the light march: the scheme took effect on the GPU
    plain = run(s, u, hydro_args(0.1, G), 5)
    assert errs(plain, ref).max() > 1e-9
    s.close()


@pytest.mark.parametrize("variant", ["ppm", "donor"])
def test_variant_fp32_vs_oracle(orc32, variant):
    nx, ny, nz = 40, 20, 12
    u = smooth_hydro(nx, ny, nz, 31).astype(np.float32)
    dtdx, g = float(np.float32(0.1)), float(np.float32(G))
    sc = HydroScheme(gamma=g, dtdx=dtdx, ppm=variant == "ppm", donor=variant == "donor")
    s = solver(nx, ny, nz, "f32")
    print(f"[hydro {variant} f32]", flush=True)
    got = run(s, u, hydro_args(dtdx, g, ppm=variant == "ppm"), 3, donor=variant == "donor")
    ref = run_oracle(orc32, u, sc, 3).astype(np.float64)
    err = errs(got, ref)
    print(f"[hydro {variant} f32]: rel-L1 {err.max():.2e}", flush=True)
    assert np.all(err <= 2e-6), err
    s.close()


@pytest.mark.parametrize("variant", ["ppm", "donor"])
@pytest.mark.parametrize("cma,prec", [(True, "f64"), (False, "f64"), (True, "f32")])
def test_variant_species_vs_oracle(orc, orc32, variant, cma, prec):
    """Two passive species (-DSCALARS, with -DCMA) under each variant.  Under PPM1D the species take
    the parabolic edges like the hydro variables (the reference's -DPPM1D -DSCALARS build leaves the
    species members of its parabolic edge states at zero, spike_25d.cu:476-481)."""
    nx, ny, nz = 33, 17, 9
    us = with_species(smooth_hydro(nx, ny, nz, 43), 44)
    o_ = orc if prec == "f64" else orc32
    cast = (lambda v: v) if prec == "f64" else (lambda v: float(np.float32(v)))
    if prec == "f32":
        us = us.astype(np.float32)
    sc = HydroScheme(gamma=cast(G), dtdx=cast(0.05), nsp=2, cma=cma, ppm=variant == "ppm",
                     donor=variant == "donor")
    s = solver(nx, ny, nz, prec)
    print(f"[hydro {variant} species {prec} cma={cma}]", flush=True)
    got = run(s, us, hydro_args(cast(0.05), cast(G), ppm=variant == "ppm"), 4, nsp=2, cma=cma,
              donor=variant == "donor")
    ref = run_oracle(o_, us, sc, 4).astype(np.float64)
    err = errs(got, ref)
    print(f"[hydro {variant} species {prec} cma={cma}]: rel-L1 {err.max():.2e}", flush=True)
    assert np.all(err <= (1e-12 if prec == "f64" else 2e-6)), err
    s.close()


@pytest.mark.parametrize("variant", ["ppm", "donor"])
def test_variant_zchunk_is_bit_invariant(variant):
    nx, ny, nz = 40, 20, 24
    u = smooth_hydro(nx, ny, nz, 3)
    s = solver(nx, ny, nz)
    outs = [run(s, u, hydro_args(0.12, G, zchunk=zc, ppm=variant == "ppm"), 3, donor=variant == "donor")
            for zc in (0, 1, 5, 24)]
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    s.close()


def test_variants_take_no_compact_format():
    s = solver(16, 16, 16, "f32")
    q = [torch.zeros((16, 16, 16), dtype=torch.float16, device="cuda") for _ in range(5)]
    o = [torch.zeros_like(x) for x in q]
    with pytest.raises(Exception):
        s.hydro_step_compact(q, o, hydro_args(0.05, G, ppm=True), "f16_prim")
    s.close()
