"""Benchmark of the GLM-MHD line-march step on one MI355X.

Workload (the turbulence configuration the metric is quoted on): 256^3 periodic box, L = 1,
gamma = 5/3, rho = 1, p = 1 (cs ~ 1.29), B0 = (0.5, 0, 0), v = 0, psi = 0; solenoidal OU forcing
in shells 1 <= |k| L / 2 pi <= 3, correlation time 0.5, Philox seed 1, fixed amplitude; PPM +
full MUSCL-Hancock + HLLD + Dedner GLM; CFL 0.3.  A "step" is one driver step of run_turb:
OU advance when due, dt from the CFL max folded into the previous step (one 8-byte read-back),
and one fused march kernel.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  value = cell-updates/s of the FP64 path (whole job, one GPU).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s + % HBM peak, 256^3 HLLD FP64 vs FP32; adjoint/forward cost"
HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector, TFLOP/s
BYTES_PER_CELL = {"f64": 144, "f32": 72}   # 9 conserved variables, one read + one write
AMP = 1.8e5                # OU amplitude (frozen): rms Mach ~1 once saturated (1.6e5 -> 0.95, 2e5 -> 1.05 at t = 1-3)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run 3 plain device copies of 256 MiB (HBM counter calibration)")
    ap.add_argument("--no-gradient", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-N", type=int, default=128)
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--no-ct", action="store_true")
    ap.add_argument("--no-hydro", action="store_true")
    ap.add_argument("--no-matrix", action="store_true")
    ap.add_argument("--hydro-N", type=int, default=480)
    ap.add_argument("--ct-N", type=int, default=512)
    ap.add_argument("--ct-steps", type=int, default=200)
    ap.add_argument("--ens-boxes", type=int, default=64)
    ap.add_argument("--zchunk", type=int, default=0,
                    help="z planes per march tile (0: autotune on the first step); fixed for profiling runs")
    ap.add_argument("--ct-zchunk", default="0",
                    help="the same for the CT march: one value, or FP64,FP32 (0: autotune)")
    ap.add_argument("--ens-config", default="",
                    help="zchunk,persistent,order for the ensemble and the device loop (empty: autotune)")
    return ap.parse_args()


ARGS = None   # the parsed command line (launch-configuration overrides)


def copy_calibration():
    """Three elementwise copies (dst = src * 1, one compute kernel each) of a 1 GiB buffer, with a
    1 GiB fill in between to push src out of the caches: a kernel of known traffic (1 GiB read,
    1 GiB written per launch) to scale the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 2 ** 30 // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    flush = torch.empty_like(src)
    for _ in range(3):
        flush.fill_(1.0)
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    return {"kernel": "elementwise mul by 1 (MulFunctor)", "bytes_read_per_launch": n * 8,
            "bytes_written_per_launch": n * 8, "launches": 3}


def make_params(N):
    from glmmhdturb_amd import Params
    return Params(N=N, gamma=5.0 / 3.0, courant=0.3, rho0=1.0, cs0=math.sqrt(5.0 / 3.0), b0=0.5,
                  turb_rms=AMP, turb_T=0.5, turb_Ndt=100, comp_frac=0.0, turb_seed=1,
                  turb_gs=64, turb_kmin=1.0, turb_kmax=3.0)


def run_gpu(N, steps, warmup, precision):
    import torch
    from glmmhdturb_amd import Solver, dxof, step_args

    p = make_params(N)
    s = Solver(p, precision)
    u1, u2 = s.new_state(), s.new_state()
    s.ic_uniform(u1, p.rho0, 1.0, (p.b0, 0.0, 0.0), p.gamma)     # rho = 1, p = 1
    f = s.setup_forcing(p, AMP)
    f.advance()
    cmax = s.ctot_max(u1)
    t = 0.0
    stream = torch.cuda.current_stream()
    ev = []

    # the launch's z-chunking is autotuned on the first step's arguments (same bits for every choice)
    if ARGS is not None and ARGS.zchunk > 0:
        s.zchunk = ARGS.zchunk
        tuned = {"zchunk": ARGS.zchunk, "fixed": True}
    else:
        tuned = s.tune_zchunk(u1, u2, step_args(p, min(p.courant * dxof(p) / cmax, f.next_time), 0.0, do_turb=True,
                                                riemann="hlld", recon="ppm", predictor="full", scheme="cube"))

    def one_step(timed):
        nonlocal u1, u2, t, cmax
        while t >= f.next_time:
            f.advance()
        dt = min(p.courant * dxof(p) / cmax, f.next_time - t)
        a = step_args(p, dt, t, do_turb=True, riemann="hlld", recon="ppm", predictor="full", scheme="cube",
                      want_cfl=True)
        if timed:
            e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        s.step(u1, u2, a)
        if timed:
            e1.record(stream)
            ev.append((e0, e1))
        u1, u2 = u2, u1
        t += dt
        cmax = s.cfl_max()

    for _ in range(warmup):
        one_step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    d = s.diagnostics(u1)
    n = N ** 3
    # after the timed steps: the march's memory floor (spike_mhd.cu -DMEMFLOOR) on the same grid,
    # state and chunking -- the kernel's loads, staging and stores with no flux computation
    a = step_args(p, 1e-3, t, do_turb=True, riemann="hlld", recon="ppm", predictor="full", scheme="cube",
                  want_cfl=False)
    for _ in range(3):
        s.march_floor(u1, u2, a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(20):
        s.march_floor(u1, u2, a)
    e1.record(stream)
    torch.cuda.synchronize()
    floor_ms = e0.elapsed_time(e1) / 20
    out = dict(value=n * steps / wall, ms_per_step=wall / steps * 1e3, kernel_ms=kern_ms, t=t, mach=d["mach"],
               finite=d["finite"], divbmax=d["divbmax"], diag=d, zchunk=tuned, floor_ms=floor_ms)
    s.close()
    return out


def run_gradient_ic(N, steps, precision):
    """Throughput of the same scheme from the spike_mhd.cu gradient IC (no uniform-state regions)."""
    import torch
    from glmmhdturb_amd import Solver, step_args

    p = make_params(N)
    s = Solver(p, precision)
    u1, u2 = s.new_state(), s.new_state()
    s.ic_gradient(u1, p.gamma)
    dt = p.courant * (1.0 / N) / s.ctot_max(u1)
    a = step_args(p, dt, 0.0, riemann="hlld", recon="ppm", predictor="full", scheme="cube", want_cfl=True)
    for _ in range(2):
        s.step(u1, u2, a); u1, u2 = u2, u1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(u1, u2, a); u1, u2 = u2, u1
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s.close()
    return N ** 3 * steps / wall


def diag_errors(d64, d32):
    """Relative error of each FP32 diagnostic against FP64.  The cross helicity sum v.B can be near
    zero, so its error is also given against the Cauchy-Schwarz bound sqrt(sum m^2/rho * sum B^2)."""
    rel = {k: abs(d32[k] - d64[k]) / max(abs(d64[k]), 1e-300) for k in ("ekin", "emag", "cross_helicity", "mass", "etot")}
    rel["cross_helicity_vs_bound"] = abs(d32["cross_helicity"] - d64["cross_helicity"]) / math.sqrt(
        d64["ke2"] * 2.0 * d64["emag"])
    rel["divbmax_abs"] = abs(d32["divbmax"] - d64["divbmax"])
    return rel


def _forced_steps(p, s, f, u1, u2, t, cmax, *, nsteps=None, t_end=None):
    """The benchmark's forced step loop (OU advance cadence, CFL dt folded into the previous step)
    for `nsteps` steps or up to `t_end`; returns the state pair, the time and the next CFL speed."""
    from glmmhdturb_amd import dxof, step_args
    n = 0
    while (nsteps is not None and n < nsteps) or (t_end is not None and t < t_end):
        while t >= f.next_time:
            f.advance()
        dt = min(p.courant * dxof(p) / cmax, f.next_time - t)
        s.step(u1, u2, step_args(p, dt, t, do_turb=True, riemann="hlld", recon="ppm", want_cfl=True))
        u1, u2 = u2, u1
        t += dt
        n += 1
        cmax = s.cfl_max()
    return u1, u2, t, cmax


def fp32_from_rest(N, steps=100):
    """FP32 against FP64 on the headline workload itself: the uniform IC at rest, the OU kick on,
    exactly `steps` forced steps in each precision (each with its own CFL dt).  Also the kinetic
    energy as a fraction of the total energy, beside FP32's relative rounding unit 2^-24: while the
    flow is that slow, the kick's energy input per step lies below the rounding unit of the FP32
    total energy and is rounded away there."""
    import torch
    from glmmhdturb_amd import Solver
    p = make_params(N)
    out = {}
    for prec in ("f64", "f32"):
        s = Solver(p, prec)
        u1, u2 = s.new_state(), s.new_state()
        s.ic_uniform(u1, p.rho0, 1.0, (p.b0, 0.0, 0.0), p.gamma)
        f = s.setup_forcing(p, AMP)
        f.advance()
        u1, u2, t, _ = _forced_steps(p, s, f, u1, u2, 0.0, s.ctot_max(u1), nsteps=steps)
        torch.cuda.synchronize()
        out[prec] = s.diagnostics(u1)
        out[prec + "_t"] = t
        s.close()
    rel = diag_errors(out["f64"], out["f32"])
    rel.update(steps=steps, t=out["f64_t"], mach=out["f64"]["mach"],
               ekin_over_etot=out["f64"]["ekin"] / out["f64"]["etot"], fp32_rounding_unit=2.0 ** -24)
    return rel


def fp32_developed(N, steps):
    """FP32 against FP64 on a developed state: the gradient IC (spike_mhd.cu:304-317) with the OU kick
    on, `steps` steps, each precision with its own CFL dt."""
    import torch
    from glmmhdturb_amd import Solver
    p = make_params(N)
    out = {}
    for prec in ("f64", "f32"):
        s = Solver(p, prec)
        u1, u2 = s.new_state(), s.new_state()
        s.ic_gradient(u1, p.gamma)
        f = s.setup_forcing(p, AMP)
        u1, u2, _, _ = _forced_steps(p, s, f, u1, u2, 0.0, s.ctot_max(u1), nsteps=steps)
        torch.cuda.synchronize()
        out[prec] = s.diagnostics(u1)
        s.close()
    return diag_errors(out["f64"], out["f32"])


def fp32_developed_turbulence(N, t_dev, steps):
    """FP32 against FP64 on the benchmark's own turbulence once it is developed: the FP64 run from
    rest to t_dev (the rms Mach number has saturated there), then `steps` steps in each precision
    from that state (rounded to FP32 for the FP32 run), the forcing advanced alike."""
    import torch
    from glmmhdturb_amd import Solver
    p = make_params(N)
    s64 = Solver(p, "f64")
    u1, u2 = s64.new_state(), s64.new_state()
    s64.ic_uniform(u1, p.rho0, 1.0, (p.b0, 0.0, 0.0), p.gamma)
    f64 = s64.setup_forcing(p, AMP)
    f64.advance()
    u1, u2, t, cmax = _forced_steps(p, s64, f64, u1, u2, 0.0, s64.ctot_max(u1), t_end=t_dev)
    d0 = s64.diagnostics(u1)
    s32 = Solver(p, "f32")
    v1, v2 = s32.new_state(), s32.new_state()
    v1.copy_(u1.float())
    f32 = s32.setup_forcing(p, AMP)
    for _ in range(f64.nsteps):
        f32.advance()
    t0 = t
    u1, u2, t, _ = _forced_steps(p, s64, f64, u1, u2, t0, cmax, nsteps=steps)
    v1, v2, _, _ = _forced_steps(p, s32, f32, v1, v2, t0, s32.ctot_max(v1), nsteps=steps)
    torch.cuda.synchronize()
    rel = diag_errors(s64.diagnostics(u1), s32.diagnostics(v1))
    rel.update(t_start=t0, mach_start=d0["mach"], steps=steps)
    s64.close()
    s32.close()
    return rel


def measured_traffic(N):
    """HBM bytes per launch of the FP64 march kernel from the committed, calibrated PMC summary."""
    path = os.path.join(ROOT, "profiles", "march_fp64_256_ou_hbm_traffic.txt")
    if N != 256 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def profiled_alu(N):
    """FP64 operations per cell-update and the VALU busy fraction of the FP64 march kernel from the
    committed PMC summaries (FLOP counts: every lane counted, FMA as 2)."""
    if N != 256:
        return None, None
    fpc = busy = None
    path = os.path.join(ROOT, "profiles", "march_256_ou_flops.txt")
    if os.path.exists(path):
        for line in open(path):
            if "FP64 FLOPs per launch" in line:
                fpc = float(line.split("=")[1].split("per cell")[0])
                break
    path = os.path.join(ROOT, "profiles", "march_fp64_256_ou_valu.txt")
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("VALU busy"):
                busy = float(line.split("=")[-1].strip().rstrip("%")) / 100.0
    return fpc, busy


def host_layout():
    """CPU count and NUMA nodes of the host the CPU legs run on."""
    nodes = []
    base = "/sys/devices/system/node"
    if os.path.isdir(base):
        nodes = sorted(d for d in os.listdir(base) if d.startswith("node") and d[4:].isdigit())
    return {"cpus": os.cpu_count(), "numa_nodes": len(nodes) or None,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}


def cpu_baseline(seconds, extra_seconds=5.0):
    """The reference's own CPU path, restated: cpu_simd.jl's chunked z-stream (cpu_simd.jl:440-601;
    oracle/cpu_stream.cpp: OpenMP over k-chunks, 8/4 lanes across x, first-touch buffers) on the
    256^3 grid, in a child process with the OpenMP binding set (OMP_PROC_BIND=close,
    OMP_PLACES=cores, all host threads).  Main figure: the GPU's scheme (full Hancock + PPM +
    HLLD, FP64); beside it the same in FP32 and the literal cpu_simd.jl scheme (PLM MonCen + LLF)
    in FP32 and FP64, plus the hydro march's oracle.  Input: the gradient IC (spike_mhd.cu:304-317),
    on which no limiter short-cuts; a fixed CFL dt; no OU kick (cpu_simd.jl has none)."""
    import subprocess
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", json.dumps([seconds, extra_seconds])]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, check=True).stdout
    res = json.loads(out.strip().splitlines()[-1])
    res["host"] = host_layout()
    res["host"].update(omp_num_threads=env["OMP_NUM_THREADS"], omp_proc_bind=env["OMP_PROC_BIND"],
                       omp_places=env["OMP_PLACES"], cpu_model=_cpu_model())
    return res


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_child(seconds, extra_seconds):
    """Runs in the child process of cpu_baseline: the z-stream legs and the hydro oracle leg."""
    import ctypes as C

    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import BUILD, Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    lib = C.CDLL(os.path.join(BUILD, "libcpustream.so"))
    lib.cs_run.restype = C.c_double
    lib.cs_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 7
    N = 256
    p = make_params(N)
    orc = Oracle("f64")
    u0 = orc.ic_gradient(N, N, N, p.gamma)
    pfl = p.smallr * p.smallc ** 2 / p.gamma
    dx = 1.0 / N
    dt = p.courant * dx / orc.ctot_max(u0, p.gamma, p.smallr, pfl)
    ch = p.courant * dx / dt / 3 * p.glm_ch_scale
    fac = math.exp(-ch * dt / (p.glm_cp_coef * p.boxlen))

    def leg(budget, prec, scheme):
        u = np.ascontiguousarray(u0.astype(np.float64 if prec == 8 else np.float32))
        el = lib.cs_run(u.ctypes.data, N, 1, prec, scheme, 0, p.gamma, p.smallr, pfl, dt, dx, ch, fac)
        steps = max(1, min(50, int(budget / max(el, 1e-3))))
        u = np.ascontiguousarray(u0.astype(np.float64 if prec == 8 else np.float32))
        el = lib.cs_run(u.ctypes.data, N, steps, prec, scheme, 0, p.gamma, p.smallr, pfl, dt, dx, ch, fac)
        assert el > 0 and np.isfinite(u).all()
        name = ("full Hancock + PPM + HLLD (the GPU benchmark's scheme)" if scheme == 1
                else "cpu_simd.jl's scheme: full Hancock + PLM MonCen + LLF")
        return dict(value=N ** 3 * steps / el, unit="cell-updates/s", cores=threads, kind="port",
                    sample=f"cpu_simd.jl chunked z-stream restated (oracle/cpu_stream.cpp), {N}^3 gradient IC, "
                           f"{'FP64' if prec == 8 else 'FP32'}, {name}, {steps} steps, {el:.1f} s, {threads} threads")

    main = leg(seconds, 8, 1)
    main["other_legs"] = {
        "fp32_same_scheme": leg(extra_seconds, 4, 1),
        "fp32_cpu_simd_scheme": leg(extra_seconds, 4, 0),
        "fp64_cpu_simd_scheme": leg(extra_seconds, 8, 0),
        "hydro_march_fp32": _cpu_hydro_leg(extra_seconds),
        "note": "fp32_cpu_simd_scheme is the configuration of the reference's published CPU figure "
                "(119 Mcell/s at 256^3 on 32 EPYC threads, REPORT.md:99-105)",
    }
    print(json.dumps(main))


def _cpu_hydro_leg(seconds, M=128):
    """The hydro march's oracle (oracle/hydro_oracle.c, OpenMP) on M^3 of the hydro leg's input
    (IC ph = 0.001 (i mod 911), dt/dx 0.05, FP32), as many steps as fit in about `seconds`."""
    import numpy as np
    from oracle import HydroScheme, Oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    orc = Oracle("f32")
    ph = (0.001 * (np.arange(M ** 3) % 911)).astype(np.float32).reshape(M, M, M)
    u = np.stack([1 + 0.3 * ph, 0.1 * ph, -0.1 * ph, 0.05 * ph, 2.5 + 0.5 * ph]).astype(np.float32)
    sc = HydroScheme(gamma=1.4, dtdx=0.05)
    steps, busy = 0, 0.0
    while busy < seconds or steps < 1:
        t0 = time.perf_counter()
        u = orc.hydro_step(u, sc)
        busy += time.perf_counter() - t0
        steps += 1
    return dict(value=M ** 3 * steps / busy, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"hydro march oracle (FP32, OpenMP, {threads} threads), {M}^3 of the hydro leg's input, "
                       f"{steps} steps, {busy:.1f} s")


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, max(0, int(math.ceil(q / 100.0 * len(xs))) - 1))]


def ens_setup(ens):
    """The ensemble's launch configuration: autotuned, or fixed by --ens-config for profiling runs."""
    if ARGS is not None and ARGS.ens_config:
        z, pm, od = (int(x) for x in ARGS.ens_config.split(","))
        ens.configure(zchunk=z, persistent=pm, order=od)
        return {"zchunk": z, "persistent": pm, "order": od, "fixed": True}
    return ens.tune(reps=3)


def run_device_loop(N, steps, warmup):
    """The headline workload driven from the device: the same box through the ensemble path with
    one box, so dt, ch, the psi decay, the ramp and the OU cadence are computed on the GPU by the
    step-control kernel, the step is a replayed HIP graph, and nothing is read back per step
    (run_turb's loop, glmmhd_turb.jl:2213-2226, without its host round trip)."""
    import torch
    from glmmhdturb_amd import BoxSpec, Ensemble, Params
    ens = Ensemble([BoxSpec(N=N)], Params(gamma=5.0 / 3.0), "f64")
    best = ens_setup(ens)
    ens.graph(True)
    ens.step(warmup)
    ens.sync()
    st = ens.stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    ens.step(steps)
    e1.record(st)
    ens.sync()
    wall = time.perf_counter() - t0
    d = ens.diagnostics(0)
    out = {"value": N ** 3 * steps / wall, "ms_per_step": wall / steps * 1e3,
           "gpu_ms_per_step": e0.elapsed_time(e1) / steps, "tuned": best, "mach": d["mach"], "finite": d["finite"],
           "note": "same workload as the headline, one-box ensemble: step control on the device, HIP graph, "
                   "no per-step read-back"}
    ens.close()
    return out


def run_ensemble(sweeps, warmup, nbox):
    """The batched-ensemble workload: nbox boxes with sizes drawn uniformly from {32,48,64,96,128}^3
    by seed, each with its own seed, forcing amplitude (rms Mach target 0.3-2.0) and B0 (beta
    0.5-100); FP64 PPM + full Hancock + HLLD + GLM; autotuned launch; the sweep captured in a HIP
    graph.  Also the same boxes stepped one at a time through glm_step with run_turb's host loop."""
    import torch
    from glmmhdturb_amd import Ensemble, Params, dxof, random_boxes, step_args
    boxes = random_boxes(nbox, seed=2024)
    ens = Ensemble(boxes, Params(gamma=5.0 / 3.0), "f64")
    best = ens_setup(ens)
    ens.graph(True)
    ens.step(warmup)
    ens.sync()
    st = ens.stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(sweeps + 1)]
    ev[0].record(st)
    for k in range(sweeps):
        ens.step(1)
        ev[k + 1].record(st)
    ens.sync()
    sweep_ms = [ev[k].elapsed_time(ev[k + 1]) for k in range(sweeps)]
    total_ms = ev[0].elapsed_time(ev[-1])
    tm = ens.timing(min(sweeps, 256))
    box_ms = [x for r in tm for x in r]
    cells = ens.cells()
    diags = [ens.diagnostics(b) for b in range(nbox)]
    # the same boxes, one glm_step launch per box and step, dt read back per box (run_turb's loop)
    seq_sweeps = 5
    ens.sync()
    solvers = ens.solvers
    cur = [ens.current(b) for b in range(nbox)]
    oth = [ens.u1[b] if cur[b] is ens.u0[b] else ens.u0[b] for b in range(nbox)]
    forc = [s.setup_forcing(s.p, b.amp) for s, b in zip(solvers, boxes)]
    cm = [s.ctot_max(u) for s, u in zip(solvers, cur)]
    tt = [0.0] * nbox
    torch.cuda.synchronize()
    t0 = None
    for k in range(seq_sweeps + 1):
        if k == 1:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        for b in range(nbox):
            s, f = solvers[b], forc[b]
            while tt[b] >= f.next_time:
                f.advance()
            dt = min(s.p.courant * dxof(s.p) / cm[b], f.next_time - tt[b])
            s.step(cur[b], oth[b], step_args(s.p, dt, tt[b], do_turb=True, riemann="hlld", recon="ppm",
                                             scheme="cube", want_cfl=True))
            cur[b], oth[b] = oth[b], cur[b]
            tt[b] += dt
            cm[b] = s.cfl_max()
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t0
    out = {
        "workload": f"{nbox} OU-driven GLM-MHD boxes, sizes uniform over {{32,48,64,96,128}}^3 (seed 2024), "
                    "per-box seed, forcing amplitude and B0; FP64 PPM + full Hancock + HLLD + GLM, CFL 0.3",
        "value": cells * sweeps / (total_ms * 1e-3), "unit": "cell-updates/s", "boxes": nbox, "cells": cells,
        "sizes": {str(n): sum(1 for b in boxes if b.N == n) for n in (32, 48, 64, 96, 128)},
        "sweeps": sweeps, "ms_per_sweep": total_ms / sweeps,
        "sweep_ms_p50": pct(sweep_ms, 50), "sweep_ms_p99": pct(sweep_ms, 99),
        "box_step_ms_p50": pct(box_ms, 50), "box_step_ms_p99": pct(box_ms, 99),
        "box_step_note": "per box and sweep: first tile start to last tile end inside the shared launch "
                         "(device wall clock), over the last min(sweeps, 256) sweeps",
        "tuned": best, "hip_graph": True,
        "finite": all(d["finite"] for d in diags),
        "mach_range": [min(d["mach"] for d in diags), max(d["mach"] for d in diags)],
        "one_box_at_a_time_value": cells * seq_sweeps / seq_s,
        "one_box_at_a_time_note": f"same boxes, glm_step per box with host dt and CFL read-back, {seq_sweeps} sweeps",
    }
    ens.close()
    return out


def ct_diag(q, dx):
    """E_kin, E_mag (cell-centred B from the face fields), cross helicity sum m.B/rho, in FP64."""
    import torch
    x = q.double()
    rho, m = x[0], x[1:4]
    bcc = torch.stack([0.5 * (x[5] + torch.roll(x[5], -1, 2)), 0.5 * (x[6] + torch.roll(x[6], -1, 1)),
                       0.5 * (x[7] + torch.roll(x[7], -1, 0))])
    dV = dx ** 3
    return {"ekin": float((0.5 * (m ** 2).sum(0) / rho).sum().item()) * dV,
            "emag": float((0.5 * (bcc ** 2).sum(0)).sum().item()) * dV,
            "cross_helicity": float(((m * bcc).sum(0) / rho).sum().item()) * dV}


def run_ct_precision(N, nsteps):
    """The precision workload on the constrained-transport path (ct_mhd.jl step!): N^3 periodic,
    rho = 1, p = 1, face-centred B0 = (0, 0, 0.2) (beta = 50), OU forcing as the turbulence
    workload, HLLD fluxes with PLM + normal-only Hancock faces; FP32 state against an FP64 run of
    the same seed after nsteps steps, each precision with its own CFL dt."""
    import torch
    from glmmhdturb_amd import Params, Solver, dxof, step_args
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3, turb_T=0.5, turb_Ndt=100, comp_frac=0.0, turb_gs=64,
               turb_kmin=1.0, turb_kmax=3.0, turb_seed=1)
    out = {}
    for prec in ("f64", "f32"):
        s = Solver(p, prec)
        q1, q2 = s.new_state(8), s.new_state(8)
        q1[0] = 1.0
        q1[4] = 1.0 / (p.gamma - 1.0) + 0.5 * 0.2 ** 2
        q1[7] = 0.2
        f = s.setup_forcing(p, AMP)
        t, cmax = 0.0, s.ct_ctot_max(q1)
        czs = [int(x) for x in (ARGS.ct_zchunk.split(",") if ARGS is not None else ["0"])]
        cz = czs[0] if prec == "f64" or len(czs) == 1 else czs[1]
        if cz > 0:
            s.ct_zchunk = cz
            tuned = {"zchunk": cz, "fixed": True}
        else:
            tuned = s.tune_zchunk(q1, q2, step_args(p, p.courant * dxof(p) / cmax, 0.0, do_turb=True, riemann="hlld",
                                                    recon="plm", predictor="oned"), ct=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            while t >= f.next_time:
                f.advance()
            dt = min(p.courant * dxof(p) / cmax, f.next_time - t)
            s.ct_step(q1, q2, step_args(p, dt, t, do_turb=True, riemann="hlld", recon="plm", predictor="oned",
                                        want_cfl=True))
            q1, q2 = q2, q1
            t += dt
            cmax = s.cfl_max()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        d = ct_diag(q1, dxof(p))
        d["divbmax"] = s.ct_divb_max(q1)
        d["value"] = N ** 3 * nsteps / wall
        d["t"] = t
        d["zchunk"] = tuned
        # GPU time of one step at the last dt without the CFL pass: the fused march kernel
        # against the multi-pass kernel sequence (ct_mhd.jl step!), 5 steps each, HIP events
        a = step_args(p, dt, t, do_turb=True, riemann="hlld", recon="plm", predictor="oned", want_cfl=False)
        st = torch.cuda.current_stream()
        for key, passes in (("march_ms", False), ("passes_ms", True)):
            s.ct_step(q1, q2, a, passes=passes)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(5):
                s.ct_step(q1, q2, a, passes=passes)
                q1, q2 = q2, q1
            e1.record(st)
            torch.cuda.synchronize()
            d[key] = e0.elapsed_time(e1) / 5
        bpc = 128 if prec == "f64" else 64
        d["march_gbs_algorithmic"] = bpc * N ** 3 / (d["march_ms"] * 1e-3) / 1e9
        out[prec] = d
        s.close()
        del q1, q2
        torch.cuda.empty_cache()
    rel = {k: abs(out["f32"][k] - out["f64"][k]) / abs(out["f64"][k]) for k in ("ekin", "emag")}
    rel["cross_helicity_vs_bound"] = abs(out["f32"]["cross_helicity"] - out["f64"]["cross_helicity"]) / math.sqrt(
        2.0 * out["f64"]["ekin"] * 2.0 * out["f64"]["emag"])
    return {"workload": f"constrained transport, {N}^3, OU-forced, B0 = (0, 0, 0.2), HLLD + PLM, {nsteps} steps",
            "fp64": out["f64"], "fp32": out["f32"], "fp32_rel_err_vs_fp64": rel,
            "divbmax_fp32": out["f32"]["divbmax"], "divbmax_fp64": out["f64"]["divbmax"],
            "note": "fused CT march (one kernel per step, plus the CFL reduction in the run loop); FP32 state "
                    "with FP64 pressure recovery and CFL; march_ms / passes_ms: GPU ms per step of the fused "
                    "march and of the multi-pass kernels (ct_mhd.jl step!) at the final dt, without the CFL"}


def run_adjoint(N, nsteps, window=50, snapshots=16):
    """The adjoint workload: N^3 FP64 PPM + full Hancock + HLLD + GLM with OU forcing (shells 1-3,
    T = 0.5), initial state rho = p = 1, B0 = (0.5, 0, 0) plus a seeded solenoidal velocity
    (shells 1-4, rms 0.1); J = magnetic energy averaged over the last `window` of `nsteps` steps;
    gradients with respect to the initial velocity and the forcing amplitude through binomial
    checkpointing with `snapshots` stored states."""
    import torch
    from glmmhdturb_amd import Params
    from glmmhdturb_amd.adjoint import (Trajectory, initial_state, recompute_steps, solenoidal_velocity,
                                        velocity_gradient)
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3, turb_T=0.5, turb_Ndt=100, comp_frac=0.0, turb_gs=64,
               turb_kmin=1.0, turb_kmax=3.0, turb_seed=1)
    tr = Trajectory(p, nsteps, window, AMP, riemann="hlld", recon="ppm", snapshots=snapshots)
    v0 = torch.from_numpy(solenoidal_velocity(N, seed=5, kmin=1.0, kmax=4.0, rms=0.1)).cuda()
    u0 = initial_state(tr.s, 1.0, 1.0, (0.5, 0.0, 0.0), p.gamma, v0)
    tr.forward(u0)                      # warm-up (kernel loads, allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = tr.forward(u0)
    fwd_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    lam0, damp = tr.pullback()
    torch.cuda.synchronize()
    adj_s = time.perf_counter() - t0
    g = velocity_gradient(lam0, 1.0, v0)
    # per-step kernel times on the recorded map
    st = torch.cuda.current_stream()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    r = tr.records[nsteps // 2]
    a, b, lam = u0.clone(), tr.s.new_state(), tr.s.new_state()
    if r.field >= 0:
        tr.s.set_afield(tr.fields[r.field])
    e0.record(st)
    for _ in range(5):
        tr.s.step(a, b, r.args)
    e1.record(st)
    for _ in range(5):
        tr.s.step_adjoint(a, b, lam, r.args, tr.fields[r.field] if r.field >= 0 else None)
    e2.record(st)
    torch.cuda.synchronize()
    step_ms, rstep_ms = e0.elapsed_time(e1) / 5, e1.elapsed_time(e2) / 5
    # central-difference spot check on one velocity direction of the recorded map
    dv = torch.from_numpy(solenoidal_velocity(N, seed=77, kmin=1.0, kmax=4.0, rms=0.1)).cuda()
    eps = 1e-6
    jp = tr.replay(initial_state(tr.s, 1.0, 1.0, (0.5, 0.0, 0.0), p.gamma, v0 + eps * dv))
    jm = tr.replay(initial_state(tr.s, 1.0, 1.0, (0.5, 0.0, 0.0), p.gamma, v0 - eps * dv))
    fd_v, ad_v = (jp - jm) / (2 * eps), float((g * dv).sum().item())
    # the workload's own protocol: 32 random solenoidal velocity directions (shells 1-4)
    errs = []
    for d in range(32):
        dv = torch.from_numpy(solenoidal_velocity(N, seed=1000 + d, kmin=1.0, kmax=4.0, rms=0.1)).cuda()
        jp = tr.replay(initial_state(tr.s, 1.0, 1.0, (0.5, 0.0, 0.0), p.gamma, v0 + eps * dv))
        jm = tr.replay(initial_state(tr.s, 1.0, 1.0, (0.5, 0.0, 0.0), p.gamma, v0 - eps * dv))
        fd, ad = (jp - jm) / (2 * eps), float((g * dv).sum().item())
        errs.append(abs(fd - ad) / abs(fd))
    errs.sort()
    out = {
        "workload": f"adjoint of {nsteps} forced steps at {N}^3, FP64 PPM + full Hancock + HLLD + GLM + OU; "
                    f"J = <E_mag> over the last {window} steps; gradients wrt the initial solenoidal velocity "
                    f"(shells 1-4, rms 0.1) and the forcing amplitude; binomial checkpointing, {snapshots} snapshots",
        "adjoint_over_forward_cost": adj_s / fwd_s, "forward_s": fwd_s, "adjoint_s": adj_s, "J": J,
        "recomputed_forward_steps": tr.forward_steps, "reverse_steps": tr.reverse_steps,
        "schedule_forward_steps": recompute_steps(nsteps, snapshots),
        "forward_step_ms": step_ms, "reverse_step_ms": rstep_ms, "reverse_over_forward_step": rstep_ms / step_ms,
        "dJ_damp": damp,
        "fd_check": {"velocity_direction": {"fd": fd_v, "adjoint": ad_v, "rel_err": abs(fd_v - ad_v) / abs(fd_v)},
                     "random_directions_32": {"within_1e-6": sum(e <= 1e-6 for e in errs), "median_rel_err": errs[16],
                                              "max_rel_err": errs[-1]},
                     "eps": eps, "map": "recorded dt, ch, psi decay, ramp and forcing realisation",
                     "note": "central differences of J; a direction whose +-eps crosses a limiter, PPM or HLLD "
                             "branch switch differs from the derivative (tests pin the adjoint to the dual-number "
                             "derivative at 16^3 and 32^3)"},
    }
    tr.close()
    # the same reverse sweep over states kept in HBM by the forward run (nothing recomputed)
    tra = Trajectory(p, nsteps, window, AMP, riemann="hlld", recon="ppm", snapshots="all")
    tra.forward(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam0a, dampa = tra.pullback()
    torch.cuda.synchronize()
    adj_a = time.perf_counter() - t0
    out["all_states_stored"] = {
        "adjoint_over_forward_cost": adj_a / fwd_s, "adjoint_s": adj_a, "recomputed_forward_steps": tra.forward_steps,
        "stored_state_GB": len(tra.states) * 9 * 8 * N ** 3 / 1e9,
        "same_gradient_bits": bool(torch.equal(lam0a, lam0)) and dampa == damp}
    tra.close()
    return out


def run_hydro(N=480, iters=30, warmup=3, dtdx=0.05):
    """The 5-variable hydro march (spike_25d.cu as build_march.sh builds it) on the reference's
    own timing protocol (spike_25d.cu:530-626): N^3 periodic, IC ph = 0.001 (i mod 911) with
    rho = 1 + 0.3 ph, m = (0.1, -0.1, 0.05) ph, E = 2.5 + 0.5 ph; dt/dx 0.05, gamma 1.4;
    3 untimed and 30 timed steps; FP32 (the reference's precision) and FP64."""
    import torch
    from glmmhdturb_amd import Params, Solver, hydro_args
    out = {"workload": f"hydro line march {N}^3 (normal-only Hancock, PLM MonCen, HLL), reference IC "
                       f"ph = 0.001 (i mod 911), dt/dx {dtdx}, {warmup} + {iters} steps"}
    for prec in ("f32", "f64"):
        s = Solver(Params(N=N, gamma=1.4), prec)
        q, o = s.new_state(5), s.new_state(5)
        i = torch.arange(N ** 3, device=s.device, dtype=torch.int64)
        ph = 0.001 * (i % 911).to(s.dtype)
        del i
        for v, (c0, c1) in enumerate(((1.0, 0.3), (0.0, 0.1), (0.0, -0.1), (0.0, 0.05), (2.5, 0.5))):
            q[v].view(-1).copy_(c0 + c1 * ph)
        del ph
        a = hydro_args(dtdx, 1.4)
        for _ in range(warmup):
            s.hydro_step(q, o, a)
            q, o = o, q
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(iters):
            s.hydro_step(q, o, a)
            q, o = o, q
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / iters
        bpc = 5 * 2 * (4 if prec == "f32" else 8)
        out[prec] = {"value": N ** 3 / (ms * 1e-3), "unit": "cell-updates/s", "ms_per_step": ms,
                     "bytes_per_cell": bpc, "achieved_GBs": bpc * N ** 3 / (ms * 1e-3) / 1e9,
                     "hbm_frac": bpc * N ** 3 / (ms * 1e-3) / HBM_PEAK, "finite": bool(torch.isfinite(q).all())}
        note(f"hydro {prec} {N}^3: {ms:.3f} ms per step")
        del q, o
        s.close()
        torch.cuda.empty_cache()
    out["fp32_over_fp64"] = out["f32"]["value"] / out["f64"]["value"]
    out["compact"] = run_hydro_compact(N, iters, warmup, dtdx)
    return out


# The reference's figures for its compact-storage builds of the hydro march (Mcell/s, RTX A6000,
# 480^3, spike_25d.cu:64-95): -DU16STATE 7410 (+11% over the fp32 build), -DF16STATE the same
# throughput, -DU16SP with 2 species 5340.
REF_COMPACT_A6000 = {"f16_prim": 7410, "u16_prim": 7410, "u16_log2_species": 5340}


def run_hydro_compact(N=480, iters=30, warmup=3, dtdx=0.05):
    """The hydro march on the compact storage formats (glm_hydro_step_compact), FP32 compute, on the
    same IC as run_hydro (primitives stored as f16 or uint16 codes; two species fractions
    0.3 + 0.2 ph and 1e-20 (1 + 0.5 ph), spike_25d.cu:546-547, as uint16 log2 codes)."""
    import numpy as np
    import torch
    from glmmhdturb_amd import Params, Solver, hydro_args
    s = Solver(Params(N=N, gamma=1.4), "f32")
    n = N ** 3
    i = torch.arange(n, device=s.device, dtype=torch.int64)
    ph = 0.001 * (i % 911).to(torch.float32)
    del i
    rho, mx, my, mz, E = 1.0 + 0.3 * ph, 0.1 * ph, -0.1 * ph, 0.05 * ph, 2.5 + 0.5 * ph
    vx, vy, vz = mx / rho, my / rho, mz / rho
    p = torch.clamp(0.4 * (E - 0.5 * rho * (vx * vx + vy * vy + vz * vz)), min=1e-6)
    shape = (N, N, N)

    def u16(x, lo, hi):      # uint16 code of x over [lo, hi] (spike_25d.cu:157-168)
        t = ((x.clamp(lo, hi) - lo) * (65535.0 / (hi - lo)) + 0.5).floor().to(torch.int32)
        return t.to(torch.int16).view(shape).contiguous()

    fmts = {
        "f16_prim": [x.to(torch.float16).view(shape).contiguous() for x in (rho, vx, vy, vz, p)],
        "u16_prim": [u16(torch.log2(rho), -32.0, 32.0), u16(vx, -8.0, 8.0), u16(vy, -8.0, 8.0), u16(vz, -8.0, 8.0),
                     u16(torch.log2(p), -32.0, 32.0)],
        "u16_log2_species": [x.view(shape).contiguous() for x in (rho, mx, my, mz, E)] +
                            [u16(torch.log2(0.3 + 0.2 * ph), -110.0, 0.0), u16(torch.log2(1e-20 * (1.0 + 0.5 * ph)), -110.0, 0.0)],
    }
    del rho, mx, my, mz, E, vx, vy, vz, p, ph
    bytes_per_cell = {"f16_prim": 20, "u16_prim": 20, "u16_log2_species": 48}
    out = {"workload": f"hydro line march {N}^3, compact storage, FP32 compute, {warmup} + {iters} steps"}
    a = hydro_args(dtdx, 1.4)
    st = torch.cuda.current_stream()
    for name, q in fmts.items():
        o = [torch.empty_like(x) for x in q]
        for _ in range(warmup):
            s.hydro_step_compact(q, o, a, name)
            q, o = o, q
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(iters):
            s.hydro_step_compact(q, o, a, name)
            q, o = o, q
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / iters
        bpc = bytes_per_cell[name]
        val = n / (ms * 1e-3)
        out[name] = {"value": val, "unit": "cell-updates/s", "ms_per_step": ms, "bytes_per_cell": bpc,
                     "achieved_GBs": bpc * n / (ms * 1e-3) / 1e9, "hbm_frac": bpc * n / (ms * 1e-3) / HBM_PEAK,
                     "reference_a6000_Mcells": REF_COMPACT_A6000[name], "ratio_vs_reference": val / 1e6 / REF_COMPACT_A6000[name]}
        note(f"hydro compact {name} {N}^3: {ms:.3f} ms per step")
        del q, o
        torch.cuda.empty_cache()
    s.close()
    return out


# The reference's march matrix: Mcell/s at 480^3 on an RTX A6000, gradient IC, FP32 state with
# an f16 tile (BENCHMARKS.md:189-195).  Keys: (march, reconstruction, species).
REF_MATRIX_A6000 = {("glm", "plm", 0): 3175, ("glm", "plm", 2): 2733, ("glm", "ppm", 0): 2064, ("glm", "ppm", 2): 1410,
                    ("ct", "plm", 0): 1255, ("ct", "plm", 2): 1011, ("ct", "ppm", 0): 752, ("ct", "ppm", 2): 558,
                    ("hydro", "plm", 0): 6865, ("hydro", "plm", 2): 5082, ("hydro", "ppm", 0): 3995,
                    ("hydro", "ppm", 2): 3822}


def run_reference_matrix(N=480, iters=30, warmup=3):
    """The reference's march matrix on its own protocol (spike_mhd.cu:343-370): N^3 periodic, the
    gradient IC ph = 0.001 (i mod 911) (spike_mhd.cu:304-317; species rho X = rho (0.2 + 0.3 ph),
    :313-315), fixed dt/dx 0.02, CH = GLMFAC = 1, 3 untimed and 30 timed steps, FP32 state.  The
    GLM and CT marches run the MarchBridge scheme (normal-only Hancock, MonCen or the parabolic-edge
    PPM, HLL, clamped faces; march_args); the hydro march the build_march.sh scheme at dt/dx 0.05
    (spike_25d.cu:180), its PPM cells the -DPPM1D parabolic edges (:467-483).  Each configuration takes the fastest z-chunk of a short sweep (the
    reference reports its best tile per configuration)."""
    import torch
    from glmmhdturb_amd import Params, Solver, hydro_args, march_args
    out = {"workload": f"reference march matrix {N}^3, gradient IC, FP32, {warmup} + {iters} steps per configuration",
           "unit": "Mcell/s", "reference": "RTX A6000, BENCHMARKS.md:189-195"}
    s = Solver(Params(N=N, gamma=1.4), "f32")
    st = torch.cuda.current_stream()

    def timed(step, zc):
        for _ in range(warmup):
            step(zc)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(st)
        for _ in range(iters):
            step(zc)
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / iters

    rows = {}
    for march, rc, nsp in sorted(REF_MATRIX_A6000):
        if march == "glm":
            nv = 9 + nsp
            q, o = s.new_state(nv), s.new_state(nv)
            s.ic_gradient(q[:9], 1.4)
        elif march == "ct":
            nv = 8 + nsp
            q, o, g = s.new_state(nv), s.new_state(nv), s.new_state(9)
            s.ic_gradient(g, 5.0 / 3.0)
            q[:8].copy_(g[:8])
            del g
        else:
            nv = 5 + nsp
            q, o = s.new_state(nv), s.new_state(nv)
        i = torch.arange(N ** 3, device=s.device, dtype=torch.int64)
        ph = (0.001 * (i % 911).to(torch.float64)).to(s.dtype)
        del i
        if march == "hydro":
            for v, (c0, c1) in enumerate(((1.0, 0.3), (0.0, 0.1), (0.0, -0.1), (0.0, 0.05), (2.5, 0.5))):
                q[v].view(-1).copy_(c0 + c1 * ph)
        for j in range(nsp):
            q[nv - nsp + j].view(-1).copy_(q[0].view(-1) * (0.2 + 0.3 * ph))
        del ph
        recon = "plm" if rc == "plm" else "ppm_spike"
        state = {"q": q, "o": o}

        def step(zc):
            if march == "glm":
                a = march_args(0.02, 1.4, recon=recon, zchunk=zc)
                if nsp:
                    s.step_species(state["q"], state["o"], a, nsp)
                else:
                    s.step(state["q"], state["o"], a)
            elif march == "ct":
                a = march_args(0.02, 5.0 / 3.0, recon=recon, zchunk=zc)
                if nsp:
                    s.ct_step_species(state["q"], state["o"], a, nsp)
                else:
                    s.ct_step(state["q"], state["o"], a)
            else:
                s.hydro_step(state["q"], state["o"], hydro_args(0.05, 1.4, zchunk=zc, ppm=rc == "ppm"), nsp=nsp,
                             cma=bool(nsp))
            state["q"], state["o"] = state["o"], state["q"]

        best = None
        for zc in (0, 16, 32, 60, 120):
            ms = timed(step, zc)
            if best is None or ms < best[1]:
                best = (zc, ms)
        mcs = N ** 3 / (best[1] * 1e-3) / 1e6
        ref = REF_MATRIX_A6000[(march, rc, nsp)]
        key = f"{march}_{rc}_{nsp}sp"
        rows[key] = {"value": mcs, "ms_per_step": best[1], "zchunk": best[0], "reference_a6000": ref,
                     "ratio_vs_reference": mcs / ref, "bytes_per_cell": 2 * 4 * nv,
                     "hbm_frac": 2 * 4 * nv * N ** 3 / (best[1] * 1e-3) / HBM_PEAK,
                     "finite": bool(torch.isfinite(state["q"]).all())}
        note(f"matrix {key}: {mcs:.0f} Mcell/s (reference A6000 {ref})")
        del q, o, state
        torch.cuda.empty_cache()
    s.close()
    out["rows"] = rows
    return out


def note(msg):
    """A progress line on stderr (the JSON record is the only line on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    global ARGS
    if len(sys.argv) > 2 and sys.argv[1] == "--cpu-child":
        cpu_child(*json.loads(sys.argv[2]))
        return
    args = parse()
    ARGS = args
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU path is "
              f"design-only)", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    N = args.N
    r64 = run_gpu(N, args.steps, args.warmup, "f64")
    note(f"fp64 {N}^3: {r64['kernel_ms']:.3f} ms per march launch")
    rec = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (uniform IC + seeded solenoidal OU forcing)",
        "config": {"workload": f"OU-driven GLM-MHD turbulence {N}^3 periodic, PPM + full Hancock + HLLD + GLM, "
                               f"CFL 0.3, FP64", "N": N, "precision": "fp64", "parallelism": "single GPU"},
    }
    bpl = BYTES_PER_CELL["f64"] * N ** 3
    ach = bpl / (r64["kernel_ms"] * 1e-3)
    rec["roofline"] = {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": ach / HBM_PEAK, "traffic": measured_traffic(N),
                       "traffic_source": "profiles/march_fp64_256_ou_hbm_traffic.txt (FETCH_SIZE x2.0 and WRITE_SIZE "
                                         "x1.0 calibrated on a 1 GiB elementwise copy)",
                       "kernel": "glm::march_kernel<double, PRED_FULL, RIEM_HLLD>",
                       "kernel_ms": r64["kernel_ms"], "bytes_per_cell": BYTES_PER_CELL["f64"]}
    # the march's own memory floor (glm_march_floor, spike_mhd.cu -DMEMFLOOR): its loads, LDS staging
    # and stores timed alone, as the algorithmic bytes per launch over that time
    rec["memory_floor"] = {"kernel": "glm::march_floor_kernel<double>", "ms": r64["floor_ms"],
                           "algorithmic_GBs": bpl / (r64["floor_ms"] * 1e-3) / 1e9,
                           "frac_of_peak": bpl / (r64["floor_ms"] * 1e-3) / HBM_PEAK,
                           "march_over_floor": r64["kernel_ms"] / r64["floor_ms"],
                           "note": "the same launch, tiling and data movement as the march with the flux "
                                   "computation left out; the march's time over it is its compute share"}
    fpc, busy = profiled_alu(N)
    if fpc is not None:
        tf = fpc * N ** 3 / (r64["kernel_ms"] * 1e-3) / 1e12
        rec["alu"] = {"flops_per_cell": fpc, "achieved_tflops": tf, "peak_tflops": FP64_PEAK_TFLOPS,
                      "frac": tf / FP64_PEAK_TFLOPS, "valu_busy": busy,
                      "source": "profiles/march_256_ou_flops.txt and profiles/march_fp64_256_ou_valu.txt "
                                "(PMC counts of the same kernel; rate from this run's kernel time)"}
    rec["state"] = {"t": r64["t"], "mach": r64["mach"], "divbmax": r64["divbmax"], "finite": r64["finite"]}
    rec["launch_tuning"] = {"fp64": r64["zchunk"],
                            "note": "z planes per march tile, autotuned on the first step (Solver.tune_zchunk); "
                                    "auto_ms is the library's occupancy model"}
    if not args.no_fp32:
        r32 = run_gpu(N, args.steps, args.warmup, "f32")
        note(f"fp32 {N}^3: {r32['kernel_ms']:.3f} ms per march launch")
        ach32 = BYTES_PER_CELL["f32"] * N ** 3 / (r32["kernel_ms"] * 1e-3)
        rel = diag_errors(r64["diag"], r32["diag"])
        rec["launch_tuning"]["fp32"] = r32["zchunk"]
        rec["fp32"] = {"value": r32["value"], "ms_per_step": r32["ms_per_step"], "kernel_ms": r32["kernel_ms"],
                       "hbm_frac": ach32 / HBM_PEAK, "speedup_vs_fp64": r32["value"] / r64["value"],
                       "steps_compared": args.steps + args.warmup,
                       "diag_rel_err_vs_fp64_from_rest": rel,
                       "diag_rel_err_vs_fp64_from_rest_100_steps": fp32_from_rest(N, 100),
                       "diag_rel_err_vs_fp64_gradient_ic_100_steps": fp32_developed(N, 100),
                       "diag_rel_err_vs_fp64_developed_turbulence_100_steps": fp32_developed_turbulence(N, 1.0, 100),
                       "note": "FP32 state and fluxes; FP64 pressure recovery, CFL max and diagnostics"}
    if not args.no_ct:
        note("fp32 diagnostics compared; constrained transport next")
        rec["ct_precision"] = run_ct_precision(args.ct_N, args.ct_steps)
    if not args.no_hydro:
        note("hydro march next")
        rec["hydro"] = run_hydro(args.hydro_N)
    if not args.no_matrix:
        note("reference march matrix next")
        rec["reference_matrix"] = run_reference_matrix(args.hydro_N)
    if not args.no_adjoint:
        note("adjoint next")
        rec["adjoint"] = run_adjoint(args.adj_N, args.adj_steps)
    if not args.no_ensemble:
        note("device-driven loop next")
        rec["device_loop"] = run_device_loop(N, args.steps, args.warmup)
        note("ensemble next")
        rec["ensemble"] = run_ensemble(args.steps, args.warmup, args.ens_boxes)
    if not args.no_gradient:
        rec["gradient_ic_fp64"] = run_gradient_ic(N, max(5, args.steps // 3), "f64")
    if args.copy_calib:
        rec["copy_calibration"] = copy_calibration()
    if not args.no_cpu:
        note("CPU baseline next")
        rec["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
