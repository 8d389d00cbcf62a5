"""Test states that drive the scheme's guarded branches (shared by the oracle and GPU tests).

``pocket_state`` is a smooth periodic MHD state with four 3x3x3 pockets cut into it.
* ``kind="void"``: rho ~ 1e-5 and p ~ 1e-9 in the pockets, each cell moving away from the
  pocket centre at 2 (a supersonic expansion).  With smallr = 1e-6 the predicted pressure drops
  to the floor (predictor revert), face pressures go negative (iflo revert), and with
  llf_dmin = 0.01 / llf_pmin = 1e-4 the HLLD faces touching a pocket switch to LLF.
* ``kind="degenerate"``: rho ~ 1e-8, p ~ 1e-14, |B| ~ 1e-7, |v| ~ 1e-3 in the pockets.  With
  smallr = 1e-12 the HLLD star denominator of faces inside a pocket is ~1e-13 (the |denK| < 1e-12
  guard), densities cross the update floor and predictor/face reverts fire.
"""
import numpy as np

POCKETS = [(3, 4, 5), (13, 9, 2), (20, 15, 11), (7, 17, 13)]
KINDS = {
    "void": dict(rho_v=1e-5, p_v=1e-9, vkick=2.0, bscale=1.0, smallr=1e-6, smallc=1e-4, llf_dmin=0.01, llf_pmin=1e-4),
    "degenerate": dict(rho_v=1e-8, p_v=1e-14, vkick=1e-3, bscale=1e-7, smallr=1e-12, smallc=1e-8, llf_dmin=0.0,
                       llf_pmin=0.0),
}


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
    bx, by, bz = 0.5 + 0.2 * np.cos(tp * Z), 0.3 * np.sin(tp * X), 0.2 * np.cos(tp * Y)
    psi = 0.01 * np.sin(tp * X)
    for ci, cj, ck in POCKETS:
        for dk in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    c = ((ck + dk) % nz, (cj + dj) % ny, (ci + di) % nx)
                    rho[c] = k["rho_v"] * (1 + 0.3 * rng.random())
                    p[c] = k["p_v"] * (1 + rng.random())
                    vx[c], vy[c], vz[c] = k["vkick"] * di, k["vkick"] * dj, k["vkick"] * dk
                    bx[c] *= k["bscale"]; by[c] *= k["bscale"]; bz[c] *= k["bscale"]
    q = np.stack([rho, vx, vy, vz, p, bx, by, bz, psi])
    u = np.empty_like(q)
    for idx in np.ndindex(nz, ny, nx):
        u[(slice(None),) + idx] = orc.prim2cons(q[(slice(None),) + idx], gamma)
    return u
