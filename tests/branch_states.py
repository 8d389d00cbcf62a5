"""Test states that drive the scheme's guarded branches (shared by the oracle and GPU tests).

``pocket_state`` is a smooth periodic MHD state with four 3x3x3 pockets cut into it.  Each kind is
chosen so that the branch fires on many cells while the run stays well conditioned: the oracle
built with and without FMA contraction agrees to ~1e-16 in relative L1 after three steps, so a
GPU-vs-oracle comparison at 1e-12 tests the branch and not the rounding.
* ``kind="void"``: rho ~ 1e-5 and p ~ 1e-9 in the pockets, each cell moving away from the
  pocket centre at 2 (a supersonic expansion).  With smallr = 1e-6 the predicted pressure drops
  to the floor (predictor revert), face pressures go negative (iflo revert), and with
  llf_dmin = 0.01 / llf_pmin = 1e-4 the HLLD faces touching a pocket switch to LLF.
* ``kind="degenerate"``: a weakly magnetised medium (|B| ~ 1e-6 everywhere) with pockets of
  rho ~ 1e-8, p ~ 1e-14, |v| ~ 1e-3.  The HLLD star denominator dL*(SL-uL)*(SL-SM) - Bn^2 of the
  faces inside a pocket is ~1e-14, under the |denK| < 1e-12 guard.  (A pocket whose field alone is
  scaled down collapses magnetically within a step and leaves a CFL-violating state whose results
  depend on rounding at the 1e-11 level.)
* ``kind="floor"``: smallr = 0.01 and pockets at rho ~ 0.0105 expanding at 2: the outflow takes
  some pocket cells below smallr, so the update's void floor (rho = smallr, m = 0, E = pfl/(gamma-1)
  + B^2/2) replaces them.
"""
import numpy as np

POCKETS = [(3, 4, 5), (13, 9, 2), (20, 15, 11), (7, 17, 13)]
# rho_v, p_v, vkick: pocket density, pressure, outward speed; ball: field scale of the whole state
KINDS = {
    "void": dict(rho_v=1e-5, p_v=1e-9, vkick=2.0, ball=1.0, smallr=1e-6, smallc=1e-4, llf_dmin=0.01, llf_pmin=1e-4),
    "degenerate": dict(rho_v=1e-8, p_v=1e-14, vkick=1e-3, ball=1e-6, smallr=1e-12, smallc=1e-8, llf_dmin=0.0,
                       llf_pmin=0.0),
    "floor": dict(rho_v=0.0105, p_v=0.01, vkick=2.0, ball=0.3, smallr=0.01, smallc=1e-4, llf_dmin=0.0, llf_pmin=0.0),
}

# GPU parity cases (tests/test_gpu_branches.py): kind, predictor, reconstruction, Riemann solver, the
# branches that must fire, and the switch that turns the first of them off
CASES = [
    ("void", "full", "ppm", "hlld", ("mh_revert", "iflo_revert", "llf_switch"), dict(llf_dmin=0.0, llf_pmin=0.0)),
    ("void", "oned", "ppm", "hlld", ("mh_revert", "iflo_revert", "llf_switch"), dict(mh_revert=False)),
    ("void", "full", "plm", "hll", ("mh_revert", "iflo_revert"), dict(face_guard="none")),
    ("void", "oned", "plm", "llf", ("iflo_revert",), dict(face_guard="none")),
    ("void", "full", "ppm_spike", "hlld", ("mh_revert", "llf_switch"), dict(mh_revert=False)),
    ("degenerate", "full", "ppm", "hlld", ("hlld_degenerate",), dict(riemann="hll")),
    ("degenerate", "oned", "plm", "hlld", ("hlld_degenerate",), dict(riemann="hll")),
    ("floor", "full", "ppm", "hlld", ("update_floor", "mh_revert", "iflo_revert"), dict(update_floor=False)),
    ("floor", "full", "plm", "hll", ("update_floor",), dict(update_floor=False)),
    ("floor", "oned", "plm", "llf", ("update_floor",), dict(update_floor=False)),
]


def pocket_state(orc, nx, ny, nz, kind="void", gamma=5.0 / 3.0, seed=7):
    k = KINDS[kind]
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    z = (np.arange(nz) + 0.5) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tp = 2 * np.pi
    rho = 1 + 0.3 * np.sin(tp * X) * np.cos(tp * Y)
    p = 0.6 + 0.2 * np.cos(tp * (Y + Z))
    vx, vy, vz = 0.5 * np.sin(tp * Y), 0.4 * np.cos(tp * Z), 0.3 * np.sin(tp * X)
    b = k["ball"]
    bx, by, bz = b * (0.5 + 0.2 * np.cos(tp * Z)), b * 0.3 * np.sin(tp * X), b * 0.2 * np.cos(tp * Y)
    psi = b * 0.01 * np.sin(tp * X)
    for ci, cj, ck in POCKETS:
        for dk in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    c = ((ck + dk) % nz, (cj + dj) % ny, (ci + di) % nx)
                    rho[c] = k["rho_v"] * (1 + 0.3 * rng.random())
                    p[c] = k["p_v"] * (1 + rng.random())
                    vx[c], vy[c], vz[c] = k["vkick"] * di, k["vkick"] * dj, k["vkick"] * dk
    q = np.stack([rho, vx, vy, vz, p, bx, by, bz, psi])
    u = np.empty_like(q)
    for idx in np.ndindex(nz, ny, nx):
        u[(slice(None),) + idx] = orc.prim2cons(q[(slice(None),) + idx], gamma)
    return u
