"""glmmhdturb_amd -- MI355X (gfx950) GLM-MHD line-march update of GLMMHDTurb.

The numerics live in libglmmhd.so (HIP kernels + C ABI, include/glmmhd.h).  This package is
the host-side mirror of the reference's Julia interfaces:

* ``march_bridge`` -- march_bridge/MarchBridge.jl (open_lib, set_dtdx, set_glm, set_gamma, run!)
* ``turb``         -- glmmhd_turb.jl's Params / step_plm! / compute_dt / diagnostics / ICs /
                      Forcing / run_ot / run_turb
* ``ensemble``     -- many run_turb boxes of uneven size advanced together (no reference
                      counterpart): Ensemble, BoxSpec, random_boxes
"""
from . import _lib  # noqa: F401
from .turb import (Params, Solver, Forcing, dxof, pfloor, glm_speed, ramp_of, step_args, run_ot,  # noqa: F401
                   run_turb, SCHEMES, hydro_args, march_args)

from .ensemble import BoxSpec, Ensemble, random_boxes, single_box_run  # noqa: F401

__all__ = ["Params", "Solver", "Forcing", "dxof", "pfloor", "glm_speed", "ramp_of", "step_args", "run_ot",
           "run_turb", "SCHEMES", "hydro_args", "march_args", "BoxSpec", "Ensemble", "random_boxes", "single_box_run"]
