"""Python mirror of march_bridge/MarchBridge.jl over libglmmhd / libmhd<N>.so / libct<N>.so.

Same names, argument meaning and error behaviour as the Julia module (MarchBridge.jl:25-81):
``open_lib`` dlopens a library and resolves ``march_nv``/``march_nx``/``march_set_dtdx``/
``march_run_dev`` (unconditionally) and ``march_set_glm`` (optionally); ``run_`` (Julia ``run!``)
takes two length-NV lists of device planes, leaves the result in ``q`` and returns the elapsed
GPU time in ms.  Planes are torch tensors on the GPU (the ROCArray of an AMDGPU.jl caller).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from ._lib import plane_array


@dataclass
class Lib:
    """MarchBridge.jl:25-32"""
    h: C.CDLL
    f_dt: object
    f_run: object
    f_glm: object | None
    NV: int
    N: int


def open_lib(path: str) -> Lib:
    """MarchBridge.jl:34-41: dlopen, read NV and N, resolve the entry points."""
    h = C.CDLL(path)
    h.march_nv.restype = C.c_int
    h.march_nx.restype = C.c_int
    NV = int(h.march_nv())
    N = int(h.march_nx())
    f_dt = h.march_set_dtdx
    f_dt.argtypes = [C.c_float]
    f_dt.restype = None
    f_run = h.march_run_dev
    f_run.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int]
    f_run.restype = C.c_double
    f_glm = getattr(h, "march_set_glm", None)   # only the GLM-MHD library exports it
    if f_glm is not None:
        f_glm.argtypes = [C.c_float, C.c_float]
        f_glm.restype = None
    return Lib(h, f_dt, f_run, f_glm, NV, N)


def close_lib(m: Lib) -> None:
    """MarchBridge.jl:43 (the library stays mapped until process exit under ctypes)."""
    m.h = None


def set_dtdx(m: Lib, v: float) -> None:
    """MarchBridge.jl:46: the kernel's fixed dt/dx."""
    m.f_dt(C.c_float(v))


def set_glm(m: Lib, ch: float, glmfac: float) -> None:
    """MarchBridge.jl:50-53: Dedner cleaning speed and psi decay factor (GLM library only)."""
    if m.f_glm is None:
        raise RuntimeError(f"this lib (NV={m.NV}) has no march_set_glm -- not a GLM-MHD build")
    m.f_glm(C.c_float(ch), C.c_float(glmfac))


def set_gamma(m: Lib, g: float) -> None:
    """MarchBridge.jl:57-61: adiabatic index."""
    f = getattr(m.h, "march_set_gamma", None)
    if f is None:
        raise RuntimeError("this lib has no march_set_gamma")
    f.argtypes = [C.c_float]
    f.restype = None
    f(C.c_float(g))


def set_grid(m: Lib, nx: int, ny: int, nz: int) -> None:
    """Runtime grid (the reference fixes it at build time, build_mhd.sh:26-28)."""
    f = m.h.march_set_grid
    f.argtypes = [C.c_int, C.c_int, C.c_int]
    f.restype = C.c_int
    if f(nx, ny, nz) != 0:
        raise RuntimeError("march_set_grid failed")
    m.N = nx


def _addrs(planes):
    """MarchBridge.jl:64: host vector of device addresses, one per SoA plane."""
    return plane_array([p.data_ptr() for p in planes])


def run_(m: Lib, q, o, nsteps: int) -> float:
    """MarchBridge.jl:73-81 (run!): nsteps periodic marches; result left in q; returns ms."""
    if not (len(q) == m.NV == len(o)):
        raise AssertionError(f"expected {m.NV} SoA planes")
    for x in q:
        if tuple(x.shape) != (m.N, m.N, m.N):
            raise AssertionError(f"planes must be {m.N}^3")
    qa, keep_q = _addrs(q)
    oa, keep_o = _addrs(o)
    ms = m.f_run(qa, oa, int(nsteps))
    if ms < 0:
        raise RuntimeError("march_run_dev failed")
    return ms


# Julia spells it run!; keep an alias with the same meaning.
run = run_
