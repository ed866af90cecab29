"""bench.py's output contract: one JSON line with the metric, the roofline of the
dominant kernel and the streaming ceiling (GPU, at a small grid so it runs in seconds),
and the refusal of more than one GPU (CPU)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=timeout)


def test_bench_refuses_more_than_one_gpu():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.gpu
def test_bench_json_line_at_64():
    print("\n[case] bench.py --n 64 --steps 20 --warmup 4 --fp64-steps 10 --no-cpu, "
          "ground-state leg at 64^3 with 64 training states", flush=True)
    r = run_bench("--n", "64", "--steps", "20", "--warmup", "4", "--fp64-steps", "10",
                  "--no-cpu", "--warm-fine", "64", "--warm-samples", "64",
                  "--warm-steps", "2000")
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    out = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "stream_ceiling", "fp64"):
        assert k in out, k
    assert out["n_gpus"] == 1 and out["steps"] == 20 and out["warmup"] == 4
    assert out["dtype"] == "f32" and out["higher_is_better"] is True
    assert out["value"] > 0 and out["ms_per_step"] > 0
    # value is grid points x steps per second (ms_per_step is printed to 5 decimals)
    assert out["value"] == pytest.approx(64 ** 3 / (out["ms_per_step"] * 1e-3) / 1e9,
                                         rel=5e-3)
    rf = out["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], abs=1e-3)
    # achieved = 80 B per point-step over the measured time
    assert rf["achieved"] == pytest.approx(64 ** 3 * 80 / (out["ms_per_step"] * 1e-3) / 1e9,
                                           rel=5e-3)
    sc = out["stream_ceiling"]
    # copy_us / add_us are printed to 0.1 us
    assert sc["step_ms"] == pytest.approx(sc["copy_us"] * 1e-3 + 2 * sc["add_us"] * 1e-3,
                                          abs=2e-4)
    assert sc["frac"] == pytest.approx(out["value"] / sc["value"], rel=1e-3)
    assert out["fp64"]["value"] > 0
    ws = out["warm_start"]
    assert "error" not in ws, ws
    assert ws["grid"] == [64, 64, 64] and ws["network"]["samples"] == 64
    assert ws["network_start"]["steps"] == 2000
    assert ws["tf_cap_steps"] == 10000 and ws["tf_start"]["steps"] <= 10000
    assert ("speedup_steps" in ws) != ("speedup_steps_lower_bound" in ws)
    cont = ws["continuation"]
    for name in ("network_start", "tf_start"):
        assert [lv["shape"][0] for lv in cont[name]["levels"]] == [21, 32, 64]
        assert cont[name]["wall_s"] > 0
    plain = cont["plain_tf_run"]
    assert plain["reached_continuation_state"] == (
        plain["energy"] <= cont["tf_start"]["energy"] and
        plain["residual"] <= cont["tf_start"]["residual"])


@pytest.mark.gpu
def test_bench_short_run_agrees_with_long_run():
    """The 16-step graphs are captured before the timed region whatever --steps and
    --warmup are: a 20-step run after 5 warm-up steps (odd parity, one graph replay and 4
    eager steps) is not slower than a 200-step run by more than 3% at 512^3."""
    ms = {}
    for steps, warm in ((20, 5), (200, 30)):
        print(f"\n[case] bench.py --n 512 --only f32 --headline-only --steps {steps} "
              f"--warmup {warm}", flush=True)
        r = run_bench("--n", "512", "--only", "f32", "--headline-only", "--no-cpu",
                      "--no-warm-start", "--steps", str(steps), "--warmup", str(warm))
        assert r.returncode == 0, r.stderr[-3000:]
        ms[steps] = json.loads(r.stdout.strip().splitlines()[-1])["ms_per_step"]
        print(f"[case] {steps} steps: {ms[steps]} ms/step", flush=True)
    assert ms[20] <= 1.03 * ms[200], ms
    assert ms[20] >= 0.85 * ms[200], ms
