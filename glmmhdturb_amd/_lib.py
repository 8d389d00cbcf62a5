"""ctypes binding of libglmmhd.so (include/glmmhd.h).

The library is built in-tree (``make -C glmmhdturb_amd``).  Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# GLMMHD_LIB selects another build of the library (e.g. a libmhd<N>.so from build_mhd.sh).
LIB_PATH = os.environ.get("GLMMHD_LIB") or os.path.join(HERE, "libglmmhd.so")

PREDICTOR = {"full": 0, "oned": 1}
RECON = {"plm": 0, "ppm": 1, "ppm_spike": 2}
SLOPE = {"moncen": 0, "minmod": 1}
RIEMANN = {"llf": 0, "llf_bound": 1, "hll": 2, "hlld": 3}
GUARD = {"none": 0, "iflo": 1, "clamp": 2}
IC = {"orszag_tang": 0, "uniform": 1, "gradient": 2}
NDIAG = 10
DIAG_KEYS = ("mass", "ke2", "divbmax", "rhomin", "finite", "ekin", "emag", "cross_helicity", "etot", "psi2")

# every symbol include/glmmhd.h declares
EXPORTS = (
    "march_nv", "march_nx", "march_nsp", "march_set_dtdx", "march_set_glm", "march_set_gamma", "march_run_dev",
    "march_set_grid", "march_run_dev_f64", "ct_run", "ct_regs", "ct_shmem",
    "glm_params_default", "glm_create", "glm_destroy", "glm_set_stream", "glm_step", "glm_step_species", "glm_march_floor", "glm_cfl_max",
    "glm_ctot_max", "glm_fast_max", "glm_diagnostics", "glm_ic", "glm_ou_setup", "glm_ou_advance", "glm_ou_afield",
    "glm_ou_steps", "glm_ou_copy_afield", "glm_ct_step", "glm_ct_step_species", "glm_ct_step_passes", "glm_ct_step_tiles", "glm_ct_ctot_max", "glm_ct_divb_max",
    "glm_hydro_step", "glm_hydro_step_scalars", "glm_hydro_step_compact", "glm_hydro_step_donor", "glm_last_error", "glm_error_string",
    "glm_kernel_info",
    "glm_ens_create", "glm_ens_destroy", "glm_ens_stream", "glm_ens_reset", "glm_ens_step", "glm_ens_graph",
    "glm_ens_sync", "glm_ens_sweeps", "glm_ens_current", "glm_ens_log", "glm_ens_timing", "glm_ens_configure",
    "glm_ens_tune",
    "glm_step_adjoint", "glm_step_adjoint_passes", "glm_adjoint_force_dot", "glm_ou_set_afield",
    "glm_diagnostics_noalloc",
)
ENS_NLOG = 8
ENS_LOG_LEN = 256
ENS_LOG_KEYS = ("sweep", "t", "dt", "ch", "glm_fac", "ramp", "cmax", "ou_due")


class GLMParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "smallr", "smallc", "pfl", "dt", "dx", "ch", "glm_fac", "llf_dmin", "llf_pmin",
        "turb_min_rho", "ramp")] + [(n, C.c_int) for n in (
        "predictor", "recon", "slope", "riemann", "mh_revert", "face_guard", "update_floor", "do_turb",
        "zchunk", "want_cfl")]


class GLMEnsBox(C.Structure):
    """glm_ens_box (include/glmmhd.h, section 3)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("do_turb", C.c_int),
                ("u0", C.c_void_p * 9), ("u1", C.c_void_p * 9),
                ("boxlen", C.c_double), ("courant", C.c_double), ("ch_scale", C.c_double), ("cp_coef", C.c_double),
                ("g", C.c_int), ("ndt", C.c_int),
                ("kmin", C.c_double), ("kmax", C.c_double), ("turb_T", C.c_double), ("comp_frac", C.c_double),
                ("amp", C.c_double), ("seed", C.c_ulonglong)]


class GLMError(RuntimeError):
    pass


_lib = None


def load(path: str | None = None):
    """Load libglmmhd.so once; raise if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise GLMError(f"{p} is missing: build it with `make -C {HERE}` (or __graft_entry__.build())")
    lib = C.CDLL(p)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    pp = C.POINTER(C.c_void_p)
    sig = {
        "march_nv": (i, []), "march_nx": (i, []), "march_nsp": (i, []),
        "march_set_dtdx": (None, [C.c_float]), "march_set_glm": (None, [C.c_float, C.c_float]),
        "march_set_gamma": (None, [C.c_float]), "march_run_dev": (d, [pp, pp, i]),
        "march_set_grid": (i, [i, i, i]), "march_run_dev_f64": (d, [pp, pp, i]),
        "ct_run": (d, [pp, pp, pp, i]), "ct_regs": (i, []), "ct_shmem": (i, []),
        "glm_params_default": (None, [C.POINTER(GLMParams)]),
        "glm_create": (vp, [i, i, i, i]), "glm_destroy": (None, [vp]), "glm_set_stream": (i, [vp, vp]),
        "glm_step": (i, [vp, pp, pp, C.POINTER(GLMParams)]),
        "glm_step_species": (i, [vp, pp, pp, C.POINTER(GLMParams), i]),
        "glm_march_floor": (i, [vp, pp, pp, C.POINTER(GLMParams)]), "glm_cfl_max": (i, [vp, C.POINTER(d)]),
        "glm_ctot_max": (i, [vp, pp, C.POINTER(GLMParams), C.POINTER(d)]),
        "glm_fast_max": (i, [vp, pp, C.POINTER(GLMParams), C.POINTER(d)]),
        "glm_diagnostics": (i, [vp, pp, C.POINTER(GLMParams), C.POINTER(d)]),
        "glm_diagnostics_noalloc": (i, [pp, i, i, i, i, C.POINTER(GLMParams), vp, vp]),
        "glm_ic": (i, [vp, pp, i, C.POINTER(d)]),
        "glm_ou_setup": (i, [vp, i, d, d, d, i, d, d, C.c_ulonglong]),
        "glm_ou_advance": (i, [vp]), "glm_ou_afield": (vp, [vp]), "glm_ou_copy_afield": (i, [vp, vp]), "glm_ou_steps": (C.c_longlong, [vp]),
        "glm_ct_step": (i, [vp, pp, pp, C.POINTER(GLMParams), i]),
        "glm_ct_step_species": (i, [vp, pp, pp, C.POINTER(GLMParams), i, i]),
        "glm_ct_step_passes": (i, [vp, pp, pp, C.POINTER(GLMParams), i]),
        "glm_ct_step_tiles": (i, [vp, pp, pp, C.POINTER(GLMParams), i]),
        "glm_ct_ctot_max": (i, [vp, pp, C.POINTER(GLMParams), C.POINTER(d)]),
        "glm_ct_divb_max": (i, [vp, pp, d, C.POINTER(d)]),
        "glm_hydro_step": (i, [vp, pp, pp, C.POINTER(GLMParams)]),
        "glm_hydro_step_scalars": (i, [vp, pp, pp, C.POINTER(GLMParams), i, i]),
        "glm_hydro_step_compact": (i, [vp, pp, pp, C.POINTER(GLMParams), i]),
        "glm_hydro_step_donor": (i, [vp, pp, pp, C.POINTER(GLMParams), i, i]),
        "glm_last_error": (i, []), "glm_error_string": (C.c_char_p, [i]),
        "glm_kernel_info": (i, [i, i, i, i, i, C.POINTER(i), C.POINTER(i)]),
        "glm_ens_create": (vp, [i, C.POINTER(GLMEnsBox), C.POINTER(GLMParams), i]),
        "glm_ens_destroy": (None, [vp]), "glm_ens_stream": (vp, [vp]), "glm_ens_reset": (i, [vp]),
        "glm_ens_step": (i, [vp, i]), "glm_ens_graph": (i, [vp, i]), "glm_ens_sync": (i, [vp]),
        "glm_ens_sweeps": (C.c_longlong, [vp]), "glm_ens_current": (i, [vp, i, pp]),
        "glm_ens_log": (i, [vp, i, i, C.POINTER(d)]), "glm_ens_timing": (i, [vp, i, C.POINTER(d)]),
        "glm_ens_configure": (i, [vp, i, i, i]), "glm_ens_tune": (i, [vp, i, C.POINTER(i), C.POINTER(d)]),
        "glm_step_adjoint": (i, [vp, pp, pp, pp, C.POINTER(GLMParams), vp]),
        "glm_step_adjoint_passes": (i, [vp, pp, pp, pp, C.POINTER(GLMParams), vp]),
        "glm_adjoint_force_dot": (i, [vp, i, C.POINTER(d)]), "glm_ou_set_afield": (i, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int) -> None:
    if code != 0:
        lib = load()
        raise GLMError(lib.glm_error_string(code).decode())


def plane_array(ptrs):
    """Host array of device plane addresses (the float** / void* const* of the ABI)."""
    arr = (C.c_void_p * len(ptrs))(*[C.c_void_p(int(p)) for p in ptrs])
    return C.cast(arr, C.POINTER(C.c_void_p)), arr
