"""CPU checks of the Julia caller side (glmmhdturb_amd/julia/): there is no Julia here, so the
files are checked against the C boundary they bind.

* Every symbol a ccall or dlsym names is declared in include/glmmhd.h and exported by the
  library it is resolved from (libglmmhd.so; the MarchBridge libraries for MarchBridge.jl).
* The Julia GlmParams struct has glm_params' fields in the header's order and C types.
* The Julia scheme tables carry the header's enum values.
* MarchBridge.jl keeps the reference module's public functions (march_bridge/MarchBridge.jl).
"""
import os
import re
import subprocess

import pytest

from test_abi_host import HEADER, header_functions

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "glmmhdturb_amd", "julia")
PKG = os.path.join(ROOT, "glmmhdturb_amd")


def jl(name):
    return open(os.path.join(JL, name)).read()


def exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def ccall_symbols(src):
    return set(re.findall(r"ccall\(\(:(\w+),\s*LIB\)", src))


def dlsym_symbols(src):
    return set(re.findall(r"(?:dlsym|_sym)\(\w+(?:\.h)?,\s*:(\w+)", src))


def test_glmmhdamd_ccalls_are_declared_and_exported():
    src = jl("GLMMHDAmd.jl")
    syms = ccall_symbols(src)
    assert {"glm_create", "glm_step", "glm_ct_step", "glm_step_adjoint", "glm_ou_setup"} <= syms
    declared = set(header_functions())
    assert syms <= declared, syms - declared
    lib = os.path.join(PKG, "libglmmhd.so")
    if not os.path.exists(lib):
        pytest.skip("libglmmhd.so not built")
    assert syms <= exports(lib), syms - exports(lib)


def test_marchbridge_symbols_exist_in_each_bridge_library():
    src = jl("MarchBridge.jl")
    syms = dlsym_symbols(src)
    assert {"march_nv", "march_nx", "march_set_dtdx", "march_run_dev", "march_set_glm", "march_set_gamma",
            "ct_run", "ct_regs", "ct_shmem", "march_set_grid", "march_run_dev_f64"} == syms
    assert syms <= set(header_functions())
    # unconditional resolutions in open_lib (dlsym without throw_error=false) must exist everywhere
    required = {"march_nv", "march_nx", "march_set_dtdx"}
    libs = [os.path.join(PKG, n) for n in ("libglmmhd.so", "libmhd480.so", "libct480.so", "libmarch480.so")]
    libs = [p for p in libs if os.path.exists(p)]
    if not libs:
        pytest.skip("bridge libraries not built")
    for p in libs:
        assert required <= exports(p), (p, required - exports(p))


def test_marchbridge_keeps_the_reference_functions():
    src = jl("MarchBridge.jl")
    for fn in ("open_lib", "close_lib", "set_dtdx", "set_glm", "set_gamma", "run!"):
        assert re.search(r"^(?:function\s+)?" + re.escape(fn) + r"\(", src, flags=re.M), fn
    assert "struct Lib" in src and "NV::Int" in src and "N::Int" in src


def test_julia_glmparams_matches_header():
    txt = open(HEADER).read()
    body = txt[txt.index("typedef struct glm_params {"):txt.index("} glm_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    want = []
    for ctype, names in re.findall(r"\b(double|int)\s+([^;]+);", body):
        for n in names.split(","):
            want.append((n.strip(), "Cdouble" if ctype == "double" else "Cint"))
    src = jl("GLMMHDAmd.jl")
    blk = src[src.index("struct GlmParams"):]
    blk = blk[:blk.index("\nend")]
    got = re.findall(r"(\w+)::(Cdouble|Cint)", blk)
    assert got == want


def test_julia_scheme_tables_match_header_enums():
    txt = open(HEADER).read()
    src = jl("GLMMHDAmd.jl")
    pairs = {"PREDICTOR": {"full": "GLM_PRED_FULL", "oned": "GLM_PRED_ONED"},
             "RECON": {"plm": "GLM_RECON_PLM", "ppm": "GLM_RECON_PPM", "ppm_spike": "GLM_RECON_PPM_SPIKE"},
             "SLOPE": {"moncen": "GLM_SLOPE_MONCEN", "minmod": "GLM_SLOPE_MINMOD"},
             "RIEMANN": {"llf": "GLM_RIEMANN_LLF", "llf_bound": "GLM_RIEMANN_LLF_BOUND", "hll": "GLM_RIEMANN_HLL",
                         "hlld": "GLM_RIEMANN_HLLD"}}
    for table, m in pairs.items():
        line = re.search(r"const " + table + r" = Dict\(([^)]*)\)", src).group(1)
        jv = dict(re.findall(r":(\w+) => (\d+)", line))
        for key, enum in m.items():
            cv = re.search(enum + r"\s*=\s*(\d+)", txt).group(1)
            assert jv[key] == cv, (table, key)
