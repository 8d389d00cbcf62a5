"""CPU tests of the boundary and the host logic: the C-ABI library loads and exports every symbol
include/glmmhd.h declares, the ctypes structs match the header, and the host-side step logic
(Dedner speed, psi decay, forcing ramp, scheme presets, Params defaults) follows the reference.
No GPU calls."""
import ctypes as C
import math
import os
import re
import subprocess

import pytest

from glmmhdturb_amd import Params, SCHEMES, _lib, dxof, glm_speed, pfloor, ramp_of, step_args

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "glmmhd.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_reference_bridge_symbols():
    names = header_functions()
    # MarchBridge.jl:34-61 resolves these by name (spike_mhd.cu:319-341); spike_ctm.cu:299-329
    for sym in ("march_nv", "march_nx", "march_nsp", "march_set_dtdx", "march_set_glm", "march_set_gamma",
                "march_run_dev", "ct_run", "ct_regs", "ct_shmem"):
        assert sym in names, sym
    assert set(names) == set(_lib.EXPORTS), set(names) ^ set(_lib.EXPORTS)


def test_library_exports_every_header_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build libglmmhd.so first (make -C glmmhdturb_amd)"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in header_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_binds_without_gpu_calls():
    lib = _lib.load()
    for name in _lib.EXPORTS:
        assert getattr(lib, name) is not None
    assert lib.march_nv() == 9 and lib.march_nsp() == 0


def test_params_struct_matches_header():
    """glm_params field order and types as declared in include/glmmhd.h."""
    txt = open(HEADER).read()
    body = txt[txt.index("typedef struct glm_params {"):txt.index("} glm_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for ctype, names in re.findall(r"\b(double|int)\s+([^;]+);", body):
        for n in names.split(","):
            fields.append((n.strip(), C.c_double if ctype == "double" else C.c_int))
    assert [(n, t) for n, t in _lib.GLMParams._fields_] == fields
    lib = _lib.load()
    p = _lib.GLMParams()
    lib.glm_params_default(C.byref(p))
    # reference defaults (glmmhd_turb.jl:28-52)
    assert p.gamma == 1.01 and p.smallr == 1e-6 and p.smallc == 1e-4 and p.turb_min_rho == 1e-5
    assert p.riemann == 3 and p.predictor == 0 and p.face_guard == 1 and p.update_floor == 1


def test_ens_box_struct_matches_header():
    """glm_ens_box field order and types as declared in include/glmmhd.h."""
    txt = open(HEADER).read()
    body = txt[txt.index("typedef struct glm_ens_box {"):txt.index("} glm_ens_box;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    ctypes_of = {"int": C.c_int, "double": C.c_double, "unsigned long long": C.c_ulonglong}
    fields = []
    for ctype, names in re.findall(r"\b(unsigned long long|double|int|void\*)\s+([^;]+);", body):
        for n in names.split(","):
            n = n.strip()
            if ctype == "void*":
                name, cnt = re.match(r"(\w+)\[(\d+)\]", n).groups()
                fields.append((name, C.c_void_p * int(cnt)))
            else:
                fields.append((n, ctypes_of[ctype]))
    got = [(n, t) for n, t in _lib.GLMEnsBox._fields_]
    assert [n for n, _ in got] == [n for n, _ in fields]
    for (n, t), (_, u) in zip(got, fields):
        assert C.sizeof(t) == C.sizeof(u) and getattr(t, "_type_", None) == getattr(u, "_type_", None), n
    assert C.sizeof(_lib.GLMEnsBox) == 248


def test_ens_create_rejects_bad_arguments_without_gpu_calls():
    lib = _lib.load()
    p = _lib.GLMParams()
    lib.glm_params_default(C.byref(p))
    boxes = (_lib.GLMEnsBox * 1)()
    assert not lib.glm_ens_create(0, boxes, C.byref(p), 8)
    assert lib.glm_last_error() == 1
    assert not lib.glm_ens_create(1, boxes, C.byref(p), 2)
    assert not lib.glm_ens_create(2000, boxes, C.byref(p), 8)
    boxes[0].nx = boxes[0].ny = boxes[0].nz = 8   # no planes, no boxlen
    assert not lib.glm_ens_create(1, boxes, C.byref(p), 8)
    assert b"bad box" in lib.glm_error_string(lib.glm_last_error())


def test_random_boxes_follow_the_workload_ranges():
    from glmmhdturb_amd import random_boxes
    from glmmhdturb_amd.ensemble import AMP_MACH1_G64, MACH_AMP_EXPONENT
    bx = random_boxes(64, seed=7)
    assert len(bx) == 64 and {b.N for b in bx} <= {32, 48, 64, 96, 128}
    assert len({b.seed for b in bx}) == 64
    for b in bx:
        beta = 2.0 / b.b0[0] ** 2
        assert 0.5 <= beta <= 100.0
        g = min(64, b.N)
        assert b.turb_gs == g
        mach = (b.amp / (AMP_MACH1_G64 * (g / 64) ** 3)) ** MACH_AMP_EXPONENT
        assert 0.3 - 1e-12 <= mach <= 2.0 + 1e-12
    assert random_boxes(64, seed=7) == bx


def test_params_defaults_follow_reference():
    p = Params()
    assert (p.N, p.boxlen, p.gamma, p.courant, p.smallr, p.smallc) == (256, 1.0, 1.01, 0.7, 1e-6, 1e-4)
    assert (p.glm_ch_scale, p.glm_cp_coef, p.turb_T, p.turb_Ndt, p.comp_frac) == (0.25, 0.18, 0.3, 100, 0.5)
    assert dxof(p) == 1.0 / 256
    assert pfloor(p) == pytest.approx(1e-6 * 1e-8 / 1.01, rel=1e-15)


def test_glm_speed_and_ramp():
    """ch = courant*dx/dt/NDIM*glm_ch_scale; glm_fac = exp(-(ch^2/(cp*L*ch))*dt) (glmmhd_turb.jl:884-886)."""
    p = Params(N=128, courant=0.4)
    dt = 1.3e-3
    ch, fac = glm_speed(p, dt)
    assert ch == pytest.approx(0.4 * (1 / 128) / dt / 3 * 0.25, rel=1e-15)
    assert fac == pytest.approx(math.exp(-ch / 0.18 * dt), rel=1e-14)
    assert glm_speed(Params(glm_cp_coef=0.0), dt)[1] == 1.0
    assert ramp_of(p, 0.15) == pytest.approx(0.5) and ramp_of(p, 10.0) == 1.0


@pytest.mark.parametrize("scheme", sorted(SCHEMES))
def test_step_args_presets(scheme):
    p = Params(N=64, gamma=5 / 3)
    a = step_args(p, 1e-3, 0.05, scheme=scheme, riemann="hlld", recon="ppm", do_turb=True)
    sc = SCHEMES[scheme]
    assert a.predictor == _lib.PREDICTOR[sc["predictor"]]
    assert a.face_guard == _lib.GUARD[sc["face_guard"]]
    assert a.mh_revert == int(sc["mh_revert"]) and a.update_floor == int(sc["update_floor"])
    if "riemann" in sc:
        assert a.riemann == _lib.RIEMANN[sc["riemann"]]
    else:
        assert a.riemann == _lib.RIEMANN["hlld"] and a.recon == _lib.RECON["ppm"]
    assert a.dt == 1e-3 and a.dx == 1 / 64 and a.do_turb == 1
    assert a.ramp == pytest.approx(0.05 / 0.3)


def test_solver_requires_gpu_and_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from glmmhdturb_amd import Solver
    with pytest.raises(_lib.GLMError):
        Solver(Params(N=8))


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.GLMError):
        _lib.load(str(tmp_path / "libglmmhd.so"))


def test_march_bridge_mirror_shapes():
    """MarchBridge.run! asserts NV and (N,N,N) shapes before any device call (MarchBridge.jl:74-75)."""
    from glmmhdturb_amd import march_bridge as MB
    m = MB.open_lib(_lib.LIB_PATH)
    assert m.NV == 9 and m.f_glm is not None

    class Fake:
        def __init__(self, shape):
            self.shape = shape

        def data_ptr(self):
            return 0

    with pytest.raises(AssertionError):
        MB.run_(m, [Fake((m.N, m.N, m.N))] * 8, [Fake((m.N, m.N, m.N))] * 8, 1)
    with pytest.raises(AssertionError):
        MB.run_(m, [Fake((4, 4, 4))] * 9, [Fake((4, 4, 4))] * 9, 1)


def test_build_scripts_reference_hipcc_for_gfx950():
    mk = open(os.path.join(ROOT, "glmmhdturb_amd", "Makefile")).read()
    assert "--offload-arch=$(ARCH)" in mk and "ARCH   ?= gfx950" in mk
    for script in ("build_mhd.sh", "build_ct.sh"):
        txt = open(os.path.join(ROOT, "glmmhdturb_amd", script)).read()
        assert "legacy" in txt and "gfx950" in txt


def test_hydro_library_exports_the_bridge_symbols():
    """libmarch480.so (build_march.sh): the symbols MarchBridge.open_lib resolves
    (MarchBridge.jl:34-41, spike_25d.cu:559-576) with march_nv = 5, no GPU calls."""
    path = os.path.join(ROOT, "glmmhdturb_amd", "libmarch480.so")
    assert os.path.exists(path), "build it with glmmhdturb_amd/build_march.sh"
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    for sym in ("march_nv", "march_nx", "march_set_dtdx", "march_run_dev", "march_set_gamma", "march_set_grid",
                "march_run_dev_f64", "glm_hydro_step"):
        assert sym in exported, sym
    lib = C.CDLL(path)
    assert lib.march_nv() == 5 and lib.march_nx() == 480


def test_species_library_exports_the_bridge_symbols():
    """libmarch480_cma.so (SCALARS=1 CMA=1 build_march.sh): march_nv = 7, march_nsp = 2."""
    path = os.path.join(ROOT, "glmmhdturb_amd", "libmarch480_cma.so")
    assert os.path.exists(path), "build it with SCALARS=1 CMA=1 glmmhdturb_amd/build_march.sh"
    lib = C.CDLL(path)
    assert lib.march_nv() == 7 and lib.march_nsp() == 2 and lib.march_nx() == 480


def _exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_glm_bridge_library_exports():
    """libmhd480.so (build_mhd.sh): the seven spike_mhd.cu symbols (spike_mhd.cu:321-340), march_nv = 9."""
    path = os.path.join(ROOT, "glmmhdturb_amd", "libmhd480.so")
    assert os.path.exists(path), "build it with glmmhdturb_amd/build_mhd.sh"
    exported = _exported(path)
    for sym in ("march_nv", "march_nx", "march_nsp", "march_set_dtdx", "march_set_glm", "march_set_gamma",
                "march_run_dev", "march_run_dev_f64"):
        assert sym in exported, sym
    lib = C.CDLL(path)
    assert lib.march_nv() == 9 and lib.march_nx() == 480 and lib.march_nsp() == 0


def test_glm_species_library_exports():
    """libmhd480_nsp2.so (NSP=2 build_mhd.sh; spike_mhd.cu built with -DNSP=2): the same symbols,
    march_nv = 11, march_nsp = 2 (spike_mhd.cu:30-33, :321-324)."""
    path = os.path.join(ROOT, "glmmhdturb_amd", "libmhd480_nsp2.so")
    assert os.path.exists(path), "build it with NSP=2 glmmhdturb_amd/build_mhd.sh"
    exported = _exported(path)
    for sym in ("march_nv", "march_nx", "march_nsp", "march_set_dtdx", "march_set_glm", "march_set_gamma",
                "march_run_dev", "march_run_dev_f64"):
        assert sym in exported, sym
    for sym in ("ct_run", "ct_regs", "ct_shmem"):
        assert sym not in exported, sym
    lib = C.CDLL(path)
    assert lib.march_nv() == 11 and lib.march_nx() == 480 and lib.march_nsp() == 2


def test_ct_library_exports_only_the_ct_bridge():
    """libct480.so (build_ct.sh): spike_ctm.cu's symbols (spike_ctm.cu:300-328) with march_nv = 8,
    and neither march_run_dev (it would read a ninth plane from the caller's 8-entry arrays) nor
    march_set_glm (MarchBridge.set_glm must raise for a non-GLM library)."""
    path = os.path.join(ROOT, "glmmhdturb_amd", "libct480.so")
    assert os.path.exists(path), "build it with glmmhdturb_amd/build_ct.sh"
    exported = _exported(path)
    for sym in ("march_nv", "march_nx", "march_nsp", "march_set_dtdx", "march_set_gamma", "ct_run", "ct_regs",
                "ct_shmem"):
        assert sym in exported, sym
    for sym in ("march_run_dev", "march_run_dev_f64", "march_set_glm"):
        assert sym not in exported, sym
    lib = C.CDLL(path)
    assert lib.march_nv() == 8 and lib.march_nsp() == 0


def test_ct_species_library_exports_only_the_ct_bridge():
    """libct480_nsp2.so (NSP=2 build_ct.sh; spike_ctm.cu built with -DNSP=2): the CT symbols with
    march_nv = 10 and march_nsp = 2 (spike_ctm.cu:301), no march_run_dev, no march_set_glm."""
    path = os.path.join(ROOT, "glmmhdturb_amd", "libct480_nsp2.so")
    assert os.path.exists(path), "build it with NSP=2 glmmhdturb_amd/build_ct.sh"
    exported = _exported(path)
    for sym in ("march_nv", "march_nx", "march_nsp", "march_set_dtdx", "march_set_gamma", "ct_run", "ct_regs",
                "ct_shmem"):
        assert sym in exported, sym
    for sym in ("march_run_dev", "march_run_dev_f64", "march_set_glm"):
        assert sym not in exported, sym
    lib = C.CDLL(path)
    assert lib.march_nv() == 10 and lib.march_nsp() == 2 and lib.march_nx() == 480


def test_hydro_libraries_have_no_glm_or_ct_entry_points():
    for name in ("libmarch480.so", "libmarch480_cma.so"):
        exported = _exported(os.path.join(ROOT, "glmmhdturb_amd", name))
        for sym in ("march_set_glm", "ct_run", "ct_regs", "ct_shmem"):
            assert sym not in exported, (name, sym)


@pytest.mark.parametrize("lib_name,nv,bridge", [("libmhd480_ppm.so", 9, "march"), ("libct480_ppm.so", 8, "ct"),
                                                ("libmarch480_ppm1d.so", 5, "march"),
                                                ("libmarch480_donor.so", 5, "march")])
def test_scheme_bridge_libraries(lib_name, nv, bridge):
    """The scheme builds (PPM=1 build_mhd.sh / build_ct.sh: -DPPM; SCHEME=ppm1d / donor
    build_march.sh: -DPPM1D / -DDONOR) export the same bridge as their default builds."""
    path = os.path.join(ROOT, "glmmhdturb_amd", lib_name)
    assert os.path.exists(path), "build it with the package's build scripts (__graft_entry__.build)"
    exported = _exported(path)
    want = ("march_run_dev", "march_run_dev_f64") if bridge == "march" else ("ct_run", "ct_regs", "ct_shmem")
    for sym in ("march_nv", "march_nx", "march_set_dtdx", "march_set_gamma") + want:
        assert sym in exported, sym
    assert ("march_set_glm" in exported) == (nv == 9)
    lib = C.CDLL(path)
    assert lib.march_nv() == nv and lib.march_nx() == 480
