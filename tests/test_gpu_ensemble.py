"""GPU tests of the ensemble (glm_ens_*, through the C ABI).

Pin: every box of an ensemble sweep is bit-identical (FP64 and FP32) to the same box run alone
through turb.Solver with run_turb's host loop.  The step scalars are checked against that host
loop too: dt, ch and ramp bit for bit, and the psi decay factor (the device exp) within 2 ulp of
math.exp; the single-box run then takes ch and psi decay from the ensemble's log so that both
marches see identical scalars.  Graph replays, launch configurations and the autotuner must not
change a single bit.
"""
import math

import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from glmmhdturb_amd import BoxSpec, Ensemble, Params, single_box_run  # noqa: E402

SCHEME = Params(gamma=5.0 / 3.0)


def boxes_small():
    """Uneven boxes (partial tiles on both axes, several z-chunks), each with its own forcing."""
    return [
        BoxSpec(N=24, b0=(0.5, 0.0, 0.0), amp=4e4, seed=11, turb_gs=24),
        BoxSpec(N=40, nx=40, ny=20, nz=36, b0=(0.2, 0.1, 0.0), amp=9e4, seed=12, turb_gs=32),
        BoxSpec(N=16, nx=16, ny=28, nz=48, b0=(1.1, 0.0, 0.3), amp=2e4, seed=13, turb_gs=16, courant=0.4),
        BoxSpec(N=32, b0=(0.05, 0.0, 0.0), do_turb=False),
        BoxSpec(N=32, nx=32, ny=36, nz=20, b0=(0.7, 0.0, 0.0), amp=1.5e5, seed=14, turb_gs=32, turb_T=0.05),
    ]


def states(ens):
    return [ens.current(b).clone() for b in range(ens.nbox)]


def ulps(a, b):
    return abs(a - b) / math.ulp(max(abs(a), abs(b)))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_boxes_bit_identical_to_single_box_runs(precision):
    nsw = 10
    ens = Ensemble(boxes_small(), SCHEME, precision)
    ens.step(nsw)
    ens.sync()
    assert ens.sweeps == nsw
    for b, spec in enumerate(ens.boxes):
        print(f"[ensemble {precision}] box {b} shape {spec.shape}", flush=True)
        log = ens.log(b, nsw)
        s, u, rec = single_box_run(spec, SCHEME, nsw, precision, log=log)
        for k in range(nsw):
            assert log[k]["sweep"] == k
            assert log[k]["t"] == rec[k]["t"] and log[k]["dt"] == rec[k]["dt"], (b, k)
            assert log[k]["ch"] == rec[k]["ch"] and log[k]["ramp"] == rec[k]["ramp"], (b, k)
            assert log[k]["cmax"] == rec[k]["cmax"], (b, k)
            fac_host = math.exp(-(rec[k]["ch"] * rec[k]["ch"] / (spec.cp_coef * spec.boxlen * rec[k]["ch"])) * rec[k]["dt"])
            assert ulps(log[k]["glm_fac"], fac_host) <= 2, (b, k)
        if spec.do_turb:
            assert sum(r["ou_due"] for r in log) >= 1
        assert torch.equal(ens.current(b), u), f"box {b} differs from its single-box run"
        s.close()
    ens.close()


def test_graph_replays_equal_eager_sweeps():
    ens = Ensemble(boxes_small(), SCHEME, "f64")
    ref = []
    for _ in range(5):
        ens.step(1)
        ens.sync()
        ref.append(states(ens))
    ens.restart()
    ens.graph(True)
    for k in range(5):
        ens.step(1)
        ens.sync()
        got = states(ens)
        # eager work of the program between replays
        torch.cuda.synchronize()
        for b in range(ens.nbox):
            assert torch.equal(got[b], ref[k][b]), (k, b)
    ens.graph(False)
    ens.close()


@pytest.mark.parametrize("cfg", [(16, 0, 0), (48, 0, 1), (0, 1, 1), (24, 1, 0), (128, 2, 1)])
def test_launch_configurations_do_not_change_results(cfg):
    ens = Ensemble(boxes_small(), SCHEME, "f64")
    ens.step(4)
    ens.sync()
    ref = states(ens)
    ens.restart()
    ens.configure(*cfg)
    ens.step(4)
    ens.sync()
    for b, r in enumerate(ref):
        assert torch.equal(ens.current(b), r), (cfg, b)
    ens.close()


def test_tuner_leaves_states_and_log_unchanged():
    ens = Ensemble(boxes_small(), SCHEME, "f64")
    ens.step(3)
    ens.sync()
    before = states(ens)
    logs = [ens.log(b, 3) for b in range(ens.nbox)]
    best = ens.tune(reps=1)
    assert best["march_ms"] > 0 and best["zchunk"] > 0 and best["persistent"] >= 0
    for b in range(ens.nbox):
        assert torch.equal(ens.current(b), before[b])
        assert ens.log(b, 3) == logs[b]
    ens.step(3)
    ens.sync()
    after = states(ens)
    ens.restart()
    ens.step(6)
    ens.sync()
    for b in range(ens.nbox):
        assert torch.equal(ens.current(b), after[b])
    ens.close()


def test_timing_records():
    ens = Ensemble(boxes_small(), SCHEME, "f64")
    ens.graph(True)
    ens.step(4)
    ens.sync()
    tm = ens.timing(4)
    assert len(tm) == 4 and all(len(r) == ens.nbox for r in tm)
    assert all(0.0 < x < 1000.0 for r in tm for x in r)
    ens.close()
