"""glm_march_floor, the memory floor of the march (spike_mhd.cu -DMEMFLOOR): glm_step's launch,
loads, staging and stores with the flux computation left out, so every cell is stored as read.
Checked bit for bit on grids with partial tiles, for several z-chunkings, in both precisions."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from glmmhdturb_amd import Params, Solver, step_args  # noqa: E402


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("zchunk", [0, 1, 7])
def test_floor_stores_the_state_it_reads(prec, zchunk):
    nx, ny, nz = 40, 28, 24                       # partial tiles in x and y
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0)
    s = Solver(p, prec)
    rng = np.random.default_rng(5)
    u = s.new_state()
    u.copy_(torch.from_numpy(rng.uniform(0.5, 2.0, (9, nz, ny, nx))).to(u.device, u.dtype))
    o = s.new_state()
    o.fill_(-1.0)
    print(f"[march floor] {prec} zchunk {zchunk}", flush=True)
    s.march_floor(u, o, step_args(p, 1e-3, 0.0, scheme="cube", riemann="hlld", recon="ppm", predictor="full",
                                  zchunk=zchunk))
    torch.cuda.synchronize()
    assert torch.equal(o, u)
    s.close()
