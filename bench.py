"""Throughput of the fused RK3 GPE stepper on one MI355X.

Workload: the 512^3 real-time GPE on [-16,16)^3 (h = 1/16), V = r^2/2, g = 500,
psi0 = TF(mu = 10) x z-line vortex + 1e-3 seeded noise (generated on the device by
libtbec's seeded generator, equal to trapbec_amd/initial.py's), dt = 5e-4, zero ghosts,
Omega = 0.
A step is one low-storage RK3 step (3 fused stage launches).  psi is resident in HBM
before the timed region; K steps are replayed as captured graphs and timed with HIP
events on the stepper's own stream, bracketed by device synchronisation.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu]

Prints one JSON line.  `value` is FP32 Gpt-steps/s with compensated stencil sums in
the centred-difference form and w rebuilt from the psi pair; FP64 and the stored-w
scheme are reported beside it (--variants adds the other FP32 summation forms).

Roofline basis: the reference's RK3 scheme moves 10 complex values per grid-point step
(80 B FP32, 160 B FP64): stage 1 reads psi, writes w and psi1; stage 2 reads psi1 and w,
writes w and psi2; stage 3 reads psi2 and w, writes psi.  `roofline.achieved` and
`frac` use that basis for every variant.  With w rebuilt (stage 2 reads psi1 and psi0,
stage 3 reads psi2 and psi1, no w) the kernels need only 8 (64 B / 128 B);
`roofline.frac_of_bytes_moved` states the fraction on that basis.

`stream_ceiling` is the same step's read/write mix moved by plain streaming kernels
(a copy and an add over n^3 complex arrays, the faster of torch's and libtbec's
hand-written non-temporal ones, timed here): the rate HBM actually delivers for that mix
on this GPU, and `stream_ceiling.frac` the stepper's value against it.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
BYTES_PER_PT_STEP = {32: 80, 64: 160}          # reference scheme, 10 transfers
MOVED_PER_PT_STEP = {(32, False): 64, (64, False): 128, (32, True): 80, (64, True): 160}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--fp64-steps", type=int, default=200)
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--seg", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["f32", "f64", "both"], default="both")
    ap.add_argument("--fp32-sum", choices=["compensated", "centred", "reference"],
                    default="centred")
    ap.add_argument("--variants", action="store_true",
                    help="also time the other FP32 summation modes")
    ap.add_argument("--headline-only", action="store_true",
                    help="time only the headline configuration of each precision (no "
                         "stored-w, rotating or tabulated-potential lines)")
    ap.add_argument("--extensions", action="store_true",
                    help="time the kernels beyond the headline path at n^3 FP32 (stage "
                         "variants per physics, potential tables, reference states, "
                         "imprint, output diagnostics) and print one JSON line")
    ap.add_argument("--imag-probe", action="store_true",
                    help="imaginary-time solve steps of the rotating 256^3 ground-state "
                         "problem (FP32 and FP64), for kernel traces")
    ap.add_argument("--continuation-probe", action="store_true",
                    help="the coarse-to-fine ground-state solve of the C5 problem from the "
                         "TF start (128^3, 192^3, 384^3), for kernel traces")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the ground-state leg (network start vs TF start)")
    ap.add_argument("--warm-fine", type=int, default=384,
                    help="grid of the ground-state leg (C5: 384^3 on [-12,12)^3)")
    ap.add_argument("--warm-steps", type=int, default=20000,
                    help="imaginary-time steps of the network start (the TF start may "
                         "take up to --warm-cap times as many)")
    ap.add_argument("--warm-cap", type=float, default=5.0)
    ap.add_argument("--warm-samples", type=int, default=2000)
    ap.add_argument("--mlp-probe", action="store_true",
                    help="for kernel traces of the warm-start network (5 -> 4 x 512 -> "
                         "98,304): forward and forward+backward at batch 32 and 256, and "
                         "a few training steps at batch 64")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: 3 FP32 stages at n^3 plus a plain "
                         "device copy of one psi-sized buffer (the counter calibration)")
    return ap.parse_args()


GRAPH_CHUNK = 16   # libtbec replays runs of steps as captured graphs of 16 steps


def warm_graphs(c, mode):
    """Capture the 16-step graphs of both buffer parities.  libtbec captures a graph
    lazily, keyed by (mode, steps, parity of the psi buffer the run starts from), the
    first time a run of that shape is replayed; a warm-up of an odd number of steps leaves
    the other parity current.  Running one chunk from each parity (an eager single step
    flips it, and flips it back) puts every capture and instantiation the timed call can
    need before the timed region, whatever --steps and --warmup are."""
    for _ in range(2):
        c.step(GRAPH_CHUNK, mode)
        c.step(1, mode)
    c.sync()


def time_steps(torch, c, steps, warmup):
    """Average ms per RK3 step over `steps` steps with HIP events on c's stream."""
    from trapbec_amd import tbec
    c.step(warmup, tbec.REAL_RK3)
    warm_graphs(c, tbec.REAL_RK3)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(c.stream_ptr())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    c.step(steps, tbec.REAL_RK3)
    e1.record(stream)
    e1.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


SUMS = {"reference": 0, "compensated": 1, "centred": 2}


def run_gpu(args, prec, steps, spec, pr, sum_name=None, w_stored=False, table=False):
    import torch
    from trapbec_amd import tbec, autotune
    sm = 0 if prec == 64 else SUMS[sum_name or args.fp32_sum]
    c = tbec.Condensate.from_problem(pr, prec=prec, sum_mode=sm)
    c.set_w_storage(w_stored)
    tile, seg, remap = autotune.lookup(pr.shape, prec, sm, w_stored, pr.Omega != 0.0)
    tile1, seg1 = autotune.lookup_first_stage(pr.shape, prec, sm, w_stored, pr.Omega != 0.0)
    if args.tile >= 0:
        tile, tile1, seg1 = args.tile, -1, 0
    if args.seg > 0:
        seg = args.seg
    c.set_tuning(tile, seg)
    c.set_first_stage_tuning(tile1, seg1)
    c.set_xcd_remap(remap)
    c.init_state(spec)            # seeded state generated in HBM (not timed)
    if table:
        # the same harmonic trap, read from a tabulated potential instead of computed:
        # the stage variant every non-harmonic kind uses (one real per point and stage)
        tab = c.potential(0.0)
        c.set_physics("table")
        c.set_potential_table(tab)
        del tab
    c.sync()
    ms, wall = time_steps(torch, c, steps, args.warmup)
    npts = pr.shape[0] * pr.shape[1] * pr.shape[2]
    gpts = npts / (ms * 1e-3) / 1e9
    achieved = npts * BYTES_PER_PT_STEP[prec] / (ms * 1e-3)
    moved_pp = MOVED_PER_PT_STEP[(prec, w_stored)] + (3 * prec // 8 if table else 0)
    moved = npts * moved_pp / (ms * 1e-3)
    c.close()
    return {"ms_per_step": ms, "gpt_steps_per_s": gpts, "wall_s": wall,
            "achieved_Bps": achieved, "moved_Bps": moved, "tile": tile, "seg": seg,
            "remap": remap, "first_stage": [tile1, seg1], "ms_per_launch": ms / 3.0,
            "w_stored": w_stored}


def stream_ceiling(n, prec, reps=20):
    """The read/write mix of one RK3 step with w rebuilt, moved by plain streaming kernels
    over n^3 complex arrays: stage 1 reads one array and writes one (a copy), stages 2
    and 3 read two and write one (an add).  copy + 2 x add is the time a step would take
    if its stages moved only their algorithmic bytes at the rate plain streaming reaches
    on this GPU.  Each operation is timed with torch's own kernels (events on torch's
    stream) and with libtbec's hand-written 16-byte non-temporal ones
    (tbec_stream_rates); the faster of the two sets the ceiling."""
    import torch
    from trapbec_amd import tbec
    dt = torch.complex64 if prec == 32 else torch.complex128
    x = torch.ones(n ** 3, dtype=dt, device="cuda:0")
    y = torch.ones_like(x)
    z = torch.empty_like(x)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    t_copy = timed(lambda: z.copy_(x))
    t_add = timed(lambda: torch.add(x, y, out=z))
    del x, y, z
    torch.cuda.empty_cache()
    esize = 8 if prec == 32 else 16
    h_copy_us, h_add_us = tbec.stream_rates(esize * n ** 3, reps)
    copy_ms = min(t_copy, h_copy_us * 1e-3)
    add_ms = min(t_add, h_add_us * 1e-3)
    step_ms = copy_ms + 2.0 * add_ms
    return {"copy_us": round(copy_ms * 1e3, 1), "add_us": round(add_ms * 1e3, 1),
            "copy_GBps": round(2 * esize * n ** 3 / (copy_ms * 1e-3) / 1e9, 1),
            "add_GBps": round(3 * esize * n ** 3 / (add_ms * 1e-3) / 1e9, 1),
            "step_ms": round(step_ms, 5),
            "value": round(n ** 3 / (step_ms * 1e-3) / 1e9, 4),
            "torch_us": {"copy": round(t_copy * 1e3, 1), "add": round(t_add * 1e3, 1)},
            "handwritten_us": {"copy": round(h_copy_us, 1), "add": round(h_add_us, 1)}}


def host_cores():
    """The host cores this job can use, and how that was read: the logical CPUs it may
    run on (sched_getaffinity), the cgroup CPU quota (cpu.max), and the physical cores
    among its CPUs (lscpu -p).  `usable` is the smallest of these."""
    info = {"nproc": os.cpu_count()}
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    info["affinity"] = len(cpus)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    info["cgroup_quota"] = quota
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True,
                             text=True, timeout=20).stdout
        cores = set()
        for ln in out.splitlines():
            if ln.startswith("#") or not ln.strip():
                continue
            cpu, core, sock = (int(v) for v in ln.split(",")[:3])
            if cpu in cpus:
                cores.add((sock, core))
        phys = len(cores) or None
        model = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for ln in model.splitlines():
            if ln.startswith("Model name:"):
                info["model"] = ln.split(":", 1)[1].strip()
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    info["physical_cores"] = phys
    usable = min(v for v in (info["affinity"], phys,
                             int(quota) if quota else None) if v)
    info["usable"] = max(1, usable)
    return info


def mpi_ranks(n, usable):
    """The rank count BASELINE.md plans: the largest power of two <= the usable cores
    whose 2-D decomposition (MPI_DIMS_CREATE) divides the n x n cross-section
    (src/mpi_var.f90:38-41 aborts otherwise)."""
    p = 1
    while 2 * p <= usable and n % (2 * p) == 0:
        p *= 2
    return p


def time_reference(n, prec, ranks, mpiexec):
    """Per-step wall time of the reference Fortran (oracle/_ref, built from its own
    sources with its Makefile's -O3) at n^3: wall(3 steps) - wall(1 step), halved, so
    start-up and initialisation cancel."""
    times = {}
    for steps in (1, 3):
        d = os.path.join(ROOT, "oracle", "_ref", f"bench{n}_f{prec}_s{steps}")
        if not os.path.exists(os.path.join(d, "run.x")):
            raise FileNotFoundError(os.path.join(d, "run.x"))
        shutil.rmtree(os.path.join(d, "data"), ignore_errors=True)
        os.makedirs(os.path.join(d, "data"))
        for r in range(ranks):
            os.makedirs(os.path.join(d, "data", f"proc{r:03d}"), exist_ok=True)
        t0 = time.perf_counter()
        res = subprocess.run([mpiexec, "-n", str(ranks), "./run.x"], cwd=d,
                             capture_output=True, text=True, errors="replace", timeout=900)
        times[steps] = time.perf_counter() - t0
        if res.returncode != 0:
            raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
        shutil.rmtree(os.path.join(d, "data"), ignore_errors=True)
        print(f"[bench] reference FP{prec} {steps} step(s) on {ranks} ranks: "
              f"{times[steps]:.2f} s", file=sys.stderr, flush=True)
    return (times[3] - times[1]) / 2.0


def cpu_baseline(n):
    """The reference Fortran on this host's usable cores, FP32 (the `value`) and FP64
    (under "fp64"), the 512^3 workload's grid, trap and dt from the reference's own
    `constant` initial state (the per-step cost does not depend on the data).  Falls
    back to the oracle's C restatement (OpenMP) if the reference build or MPI is
    absent."""
    host = host_cores()
    mpiexec = "/opt/conda/bin/mpiexec"
    ranks = mpi_ranks(n, host["usable"])
    have = all(os.path.exists(os.path.join(ROOT, "oracle", "_ref", f"bench{n}_f{p}_s{s}",
                                           "run.x")) for p in (32, 64) for s in (1, 3))
    if have and shutil.which(mpiexec):
        per = {p: time_reference(n, p, ranks, mpiexec) for p in (32, 64)}
        dims = f"{ranks} ranks"
        out = {"value": n ** 3 / per[32] / 1e9, "unit": "Gpt-steps/s", "cores": ranks,
               "kind": "reference",
               "sample": f"reference Fortran FP32 {n}^3 (amdflang -O3, MPICH), mpiexec -n "
                         f"{ranks}: wall(3 steps) - wall(1 step) = {2 * per[32]:.2f} s for "
                         f"2 steps",
               "s_per_step": per[32],
               "fp64": {"value": n ** 3 / per[64] / 1e9, "unit": "Gpt-steps/s",
                        "cores": ranks, "s_per_step": per[64],
                        "sample": f"reference Fortran FP64 (-fdefault-real-8) {n}^3, "
                                  f"{dims}, same method"},
               "host": host}
        return out
    # fallback: the C restatement with OpenMP threads, 1 step at n^3
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import numpy as np
    cores = host["usable"]
    os.environ["OMP_NUM_THREADS"] = str(cores)
    p = oracle.params((n, n, n), 32.0, (1024.0,) * 3, c0=0.0, gnl=1000.0, dt=5e-4)
    s = oracle.Stepper(p, np.full((n, n, n), 0.1 + 0.05j, np.complex64), prec=32)
    t0 = time.perf_counter()
    s.rk3(1)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 / dt / 1e9, "unit": "Gpt-steps/s", "cores": cores,
            "kind": "port", "sample": f"oracle C restatement FP32 {n}^3, 1 step, "
                                      f"{cores} OpenMP threads", "s_per_step": dt,
            "host": host}


WARM_TEST = (1.0, 0.8, 3000.0, 0.5, 0.0)   # w_perp, w_z, g, Omega, log N (held out)


def warm_start_leg(args):
    """The ground-state half of the metric (C5): the warm-start network (5 -> 4 x 512 GELU
    -> 32^3 x 3, BF16 MFMA) trained on the GPU on ground states from the batched solver,
    then, at a held-out rotating problem on args.warm_fine^3, the network's start against
    the plain TF start at equal final energy and GP residual: the network start runs
    args.warm_steps FP32 imaginary-time steps (dtau = 8e-4), and the TF start runs until
    it reaches that energy and residual, or args.warm_cap times the steps.  Then the
    coarse-to-fine solve to the gated stop rule from both starts (`continuation`), with
    the plain TF run as a lower bound on one-grid imaginary time to the same state."""
    from trapbec_amd import warmstart as ws
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    m, scale, info = ws.train_network(samples=args.warm_samples, log=log)
    try:
        coarse, infer_s = ws.predict(m, scale, WARM_TEST)
    finally:
        m.close()
    dtau = 8e-4
    pr = ws.fine_problem(WARM_TEST, args.warm_fine, dtau)
    r = ws.equal_state_comparison(pr, coarse, args.warm_steps, args.warm_cap,
                                  tuning=ws.C5_TUNING, infer_s=infer_s, log=log)
    r.pop("records")
    # the coarse-to-fine solve (128^3 -> 192^3 -> fine, gated stop rule 1e-10 with residual
    # <= 5e-4 at the fine grid) from the network's field and from the TF start; the plain
    # TF run above bounds how long one-grid imaginary time takes to reach that state
    cont = {}
    for name, field in (("network_start", coarse), ("tf_start", None)):
        cont[name] = ws.continuation_solve(pr, coarse=field, tuning=ws.C5_TUNING, log=log)
        cont[name].pop("fine_steps")
    cont["network_start"]["wall_s"] += infer_s
    plain = r["tf_start"]
    cs = cont["tf_start"]
    reached = plain["energy"] <= cs["energy"] and plain["residual"] <= cs["residual"]
    cont["plain_tf_run"] = {"steps": plain["steps"], "wall_s": plain["wall_s"],
                            "energy": plain["energy"], "residual": plain["residual"],
                            "reached_continuation_state": reached}
    if not reached:
        for name in ("network_start", "tf_start"):
            cont[f"speedup_wall_lower_bound_{name}"] = plain["wall_s"] / cont[name]["wall_s"]
    log(f"[bench] continuation: network start {cont['network_start']['wall_s']:.1f} s, TF "
        f"start {cs['wall_s']:.1f} s (E {cs['energy']:.10f} residual {cs['residual']:.2e}); "
        f"plain TF after {plain['wall_s']:.1f} s: E {plain['energy']:.10f} residual "
        f"{plain['residual']:.2e}")
    out = {"metric": "ground-state time, network (warm) start vs TF start, at equal final "
                     "energy and GP residual",
           "test_params": {"w_perp": WARM_TEST[0], "w_z": WARM_TEST[1], "g": WARM_TEST[2],
                           "Omega": WARM_TEST[3], "log_N": WARM_TEST[4]},
           "grid": [args.warm_fine] * 3, "box": [-12.0, 12.0], "dtau": dtau,
           "dtype": "f32 (FP64 reductions)", "network": info, "inference_s": infer_s, **r,
           "continuation": cont}
    sp = r.get("speedup_steps")
    log(f"[bench] warm start: network {r['network_start']['steps']} steps "
        f"{r['network_start']['wall_s']:.1f} s; TF "
        f"{'reached' if r['tf_start']['reached'] else 'did not reach'} the target in "
        f"{r['tf_start']['steps']} steps {r['tf_start']['wall_s']:.1f} s; speed-up "
        f"{sp if sp else '>= %.2f' % r['speedup_steps_lower_bound']}")
    return out


def mlp_probe(args):
    """The warm-start network's kernels at production size, for kernel traces: 5 forward
    passes and 5 loss/gradient passes at batch 32 and at batch 256, then one epoch of
    training at batch 64 over 256 samples (Adam and the BF16 weight refresh).  Prints host
    wall times per call (they include the host copies of the outputs)."""
    import numpy as np
    from trapbec_amd import tbec
    rng = np.random.default_rng(1)
    n_out = 3 * 32 ** 3
    res = {"mlp_probe": "5 -> 4 x 512 GELU -> 98304"}
    with tbec.MLP(5, 512, 4, n_out, seed=1) as m:
        for B in (32, 256):
            X = rng.uniform(-1, 1, (B, 5)).astype(np.float32)
            T = (0.3 * rng.standard_normal((B, n_out))).astype(np.float32)
            m.forward(X)
            t0 = time.perf_counter()
            for _ in range(5):
                m.forward(X)
            res[f"forward_b{B}_ms"] = round((time.perf_counter() - t0) * 200, 3)
            t0 = time.perf_counter()
            for _ in range(5):
                m.loss_grad(X, T, with_grads=False)
            res[f"loss_grad_b{B}_ms"] = round((time.perf_counter() - t0) * 200, 3)
        X = rng.uniform(-1, 1, (256, 5)).astype(np.float32)
        T = (0.3 * rng.standard_normal((256, n_out))).astype(np.float32)
        m.set_data(X, T)
        t0 = time.perf_counter()
        m.train(1, 64, 1e-3, seed=2)
        res["train_step_b64_ms"] = round((time.perf_counter() - t0) * 1e3 / 4, 3)
    print(json.dumps(res), flush=True)


def pmc_probe(args):
    """Kernels for counter runs: a device copy of n^3 complex64 (reads and writes 8 n^3
    bytes: the counter calibration) and two real-time steps (6 stage launches) of the
    FP32 stepper on a constant state, with the headline settings."""
    import numpy as np
    import torch
    from trapbec_amd import initial, tbec, autotune
    n = args.n
    src = torch.ones(n ** 3, dtype=torch.complex64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)          # reads and writes n^3 * 8 bytes
    torch.cuda.synchronize()
    del src, dst
    pr = initial.Problem((n, n, n), (32.0, 32.0, 32.0), (1.0, 1.0, 1.0), 500.0, 0.0, 5e-4)
    c = tbec.Condensate.from_problem(pr, prec=32, sum_mode=SUMS[args.fp32_sum])
    tile, seg, remap = autotune.lookup(pr.shape, 32, SUMS[args.fp32_sum])
    c.set_tuning(tile, seg)
    c.set_first_stage_tuning(*autotune.lookup_first_stage(pr.shape, 32, SUMS[args.fp32_sum]))
    c.set_xcd_remap(remap)
    psi = torch.full((n ** 3,), 0.1 + 0.05j, dtype=torch.complex64, device="cuda:0")
    c.upload_device(psi.data_ptr())
    c.sync()
    c.step(2, tbec.REAL_RK3)
    c.sync()
    print(json.dumps({"pmc_probe": n, "copy_bytes": n ** 3 * 8,
                      "stage_bytes_reference_scheme_avg": n ** 3 * 80 / 3,
                      "stage_bytes_moved_avg": n ** 3 * 64 / 3}), flush=True)


def continuation_probe(args):
    """The coarse-to-fine ground-state solve of C5's problem (384^3 rotating, dtau 8e-4)
    from the TF start, for kernel traces: levels 128^3 and 192^3, then 384^3, each to the
    gated stop rule.  Prints the per-level steps and wall times."""
    import numpy as np
    from trapbec_amd import warmstart as ws
    pr = ws.fine_problem(np.asarray(WARM_TEST), 384, 8e-4)
    r = ws.continuation_solve(pr, tuning=ws.C5_TUNING,
                              log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps({"continuation_probe": "C5 problem, TF start", **r}), flush=True)


def imag_probe(args):
    """Imaginary-time solve steps of the rotating 256^3 ground-state problem
    (initial.config_c3: Omega = 0.7, dtau = 1e-3), FP32 then FP64, each with the tile
    recorded for it: 200 warm-up steps, then 1000 steps checked every 200 (a convergence
    tolerance of 1e-30 never stops them).  Prints ms per step on the host clock."""
    import torch
    from trapbec_amd import initial, tbec
    torch.cuda.synchronize()
    pr, psi0 = initial.config_c3(n=256)
    out = {"imag_probe": "256^3 rotating imaginary time"}
    for prec, (tile, seg, remap) in ((32, (2, 32, 1)), (64, (0, 64, 0))):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.set_tuning(tile, seg)
            c.set_xcd_remap(remap)
            c.upload(psi0)
            c.imag_solve(200, 1e-30, 1.0, 200)
            c.sync()
            t0 = time.perf_counter()
            c.imag_solve(1000, 1e-30, 1.0, 200)
            c.sync()
            out[f"fp{prec}_ms_per_step"] = round((time.perf_counter() - t0) * 1e3 / 1000, 5)
            out[f"fp{prec}_tile"] = [tile, seg, remap]
    print(json.dumps(out), flush=True)


def extensions_probe(args):
    """ms per step of the stage variants each physics choice selects, and the one-off
    kernels (potential tables, the mobius Trap at (n/2)^3, reference states, the vortex
    imprint, velocity / quantum pressure / vortex cores with their host copies) at n^3
    FP32, centred sums, w rebuilt, tuned tiles.  Step timings use HIP events on the
    handle's stream; calls that synchronise are timed on the host clock."""
    import torch
    from trapbec_amd import initial, tbec, autotune
    n = args.n
    pr, spec = initial.ic_spec("m", n)
    out = {"extensions": f"{n}^3 FP32"}

    def handle():
        c = tbec.Condensate.from_problem(pr, prec=32, sum_mode=2)
        tile, seg, remap = autotune.lookup(pr.shape, 32, 2, False, False)
        c.set_tuning(tile, seg)
        c.set_xcd_remap(remap)
        c.init_state(spec)
        return c

    def step_ms(c, steps=40, warm=4):
        c.step(warm)
        warm_graphs(c, tbec.REAL_RK3)
        stream = torch.cuda.ExternalStream(c.stream_ptr())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        c.step(steps)
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / steps

    def wall(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    def table(c):
        t = c.potential(0.0)
        c.set_physics("table")
        c.set_potential_table(t)

    variants = [("harmonic_analytic", lambda c: None),
                ("harmonic_as_table", table),
                ("channel_static_table", lambda c: c.set_physics("channel")),
                ("harmonic_advective", lambda c: c.set_physics("harmonic", advective_vx=0.3)),
                ("harmonic_ring_closing", lambda c: (c.set_physics("harmonic_ring"),
                                                     setattr(c, "t", 49.96))),
                ("box_trap_refresh_every_step", lambda c: c.set_physics("box_trap"))]
    steps = {}
    for name, setup in variants:
        c = handle()
        setup(c)
        steps[name] = round(step_ms(c), 4)
        print(f"[ext] {name}: {steps[name]} ms/step", file=sys.stderr, flush=True)
        c.close()
    out["ms_per_step"] = steps
    once = {}
    c = handle()
    for kind in ("constant", "lattice", "vortex_rings", "random_phase"):
        once[f"reference_state_{kind}"] = round(wall(lambda: c.reference_state(
            kind, vor_ring_count=3, seed=1)), 3)
    c.init_state(spec)
    stream = torch.cuda.ExternalStream(c.stream_ptr())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    c.imprint_vortex(1.0, -2.0, 1.0)
    e1.record(stream)
    e1.synchronize()
    once["imprint_vortex"] = round(e0.elapsed_time(e1), 3)
    once["potential_table_to_host"] = round(wall(lambda: c.potential(0.0)), 3)
    once["velocity_with_host_copy"] = round(wall(c.velocity), 3)
    once["qpressure_with_host_copy"] = round(wall(c.qpressure), 3)
    once["vortex_cores_z"] = round(wall(lambda: c.vortex_cores(1e-5 * 10.0 / 500.0)), 3)
    c.close()
    h = n // 2
    pr2 = initial.Problem((h, h, h), (64.0, 64.0, 16.0), (1.0, 1.0, 1.0), 500.0, 0.0, 5e-4)
    c = tbec.Condensate.from_problem(pr2, prec=32, sum_mode=2)
    once[f"mobius_trap_{h}^3"] = round(wall(lambda: c.set_physics("mobius")), 3)
    c.close()
    out["ms_once"] = once
    for k, v in once.items():
        print(f"[ext] {k}: {v} ms", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the stepper "
              "runs one condensate per GPU; multi-GPU sharding is a design note only)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from trapbec_amd import initial

    n = args.n
    if args.pmc_probe:
        pmc_probe(args)
        return
    if args.mlp_probe:
        mlp_probe(args)
        return
    if args.extensions:
        extensions_probe(args)
        return
    if args.imag_probe:
        imag_probe(args)
        return
    if args.continuation_probe:
        continuation_probe(args)
        return
    pr, spec = initial.ic_spec("m", n)
    res = {}
    for prec in (32, 64):
        if args.only != "both" and args.only != f"f{prec}":
            continue
        steps = args.steps if prec == 32 else args.fp64_steps
        print(f"[bench] fp{prec}: {steps} timed steps", file=sys.stderr, flush=True)
        res[prec] = run_gpu(args, prec, steps, spec, pr)
        print(f"[bench] fp{prec}: {res[prec]['ms_per_step']:.3f} ms/step "
              f"{res[prec]['gpt_steps_per_s']:.2f} Gpt-steps/s", file=sys.stderr, flush=True)
        if not args.headline_only:
            v = run_gpu(args, prec, steps, spec, pr, None, True)
            res.setdefault("stored", {})[prec] = v
            print(f"[bench] fp{prec} stored w: {v['ms_per_step']:.3f} ms/step",
                  file=sys.stderr, flush=True)
        if prec == 32 and not args.headline_only:
            # the same workload in a frame rotating at Omega = 0.7 (adds the D1 terms)
            import dataclasses
            v = run_gpu(args, prec, steps, spec, dataclasses.replace(pr, Omega=0.7))
            res["rot"] = v
            print(f"[bench] fp32 Omega=0.7: {v['ms_per_step']:.3f} ms/step", file=sys.stderr,
                  flush=True)
            v = run_gpu(args, prec, steps, spec, pr, table=True)
            res["table"] = v
            print(f"[bench] fp32 tabulated potential: {v['ms_per_step']:.3f} ms/step",
                  file=sys.stderr, flush=True)
        if prec == 32 and args.variants:
            for name in SUMS:
                if name == args.fp32_sum:
                    continue
                v = run_gpu(args, 32, steps, spec, pr, name)
                res.setdefault("variants", {})[name] = v
                print(f"[bench] fp32 {name}: {v['ms_per_step']:.3f} ms/step", file=sys.stderr,
                      flush=True)
        torch.cuda.empty_cache()
        res.setdefault("stream", {})[prec] = stream_ceiling(n, prec)
        print(f"[bench] fp{prec} stream ceiling: {res['stream'][prec]['step_ms']:.3f} "
              "ms/step", file=sys.stderr, flush=True)
    head = 32 if 32 in res else 64
    r = res[head]
    out = {
        "metric": "Gpt-steps/s at 512\u00b3 FP32/FP64 (% HBM peak); ground-state time warm "
                  "vs TF start",
        "value": round(r["gpt_steps_per_s"], 4),
        "unit": "Gpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps if head == 32 else args.fp64_steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == 32 else "f64",
        "data": "synthetic (seeded TF + vortex + noise initial state)",
        "config": {"workload": f"real-time GPE RK3, {n}^3, box [-16,16)^3, g=500, dt=5e-4, "
                               f"zero ghosts, Omega=0",
                   "grid": [n, n, n], "stencil": "6th-order FD, fused RK3 stage",
                   "fp32_sum": args.fp32_sum, "w": "rebuilt", "tile": r["tile"],
                   "seg": r["seg"], "xcd_remap": r["remap"],
                   "first_stage_tile_seg": r["first_stage"]},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_Bps"] / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(r["achieved_Bps"] / HBM_PEAK, 4),
                     "traffic": None,
                     "kernel": "gpe_stage (one RK3 stage)",
                     "basis": "reference RK3 scheme: 10 complex transfers per point-step",
                     "bytes_per_launch": n ** 3 * BYTES_PER_PT_STEP[head] / 3,
                     "ms_per_launch": round(r["ms_per_launch"], 5),
                     "bytes_moved_per_launch": n ** 3 * MOVED_PER_PT_STEP[(head, False)] / 3,
                     "frac_of_bytes_moved": round(r["moved_Bps"] / HBM_PEAK, 4)},
    }

    def brief(v, steps):
        return {"value": round(v["gpt_steps_per_s"], 4), "unit": "Gpt-steps/s",
                "ms_per_step": round(v["ms_per_step"], 5),
                "frac": round(v["achieved_Bps"] / HBM_PEAK, 4),
                "frac_of_bytes_moved": round(v["moved_Bps"] / HBM_PEAK, 4),
                "achieved_GBps": round(v["achieved_Bps"] / 1e9, 1),
                "tile": v["tile"], "seg": v["seg"], "first_stage_tile_seg": v["first_stage"],
                "steps": steps}
    sc = res["stream"][head]
    out["stream_ceiling"] = dict(sc, frac=round(out["value"] / sc["value"], 4),
                                 basis="copy (stage 1) + 2 x add (stages 2, 3) over "
                                       f"{n}^3 complex arrays, the faster of torch's and "
                                       "hand-written non-temporal kernels for each")
    if 64 in res and head == 32:
        out["fp64"] = brief(res[64], args.fp64_steps)
        out["fp64"]["stream_ceiling"] = dict(res["stream"][64], frac=round(
            out["fp64"]["value"] / res["stream"][64]["value"], 4))
    for prec, v in res.get("stored", {}).items():
        out[f"fp{prec}_stored_w"] = brief(v, args.steps if prec == 32 else args.fp64_steps)
    if "rot" in res:
        out["fp32_rotating_omega_0.7"] = brief(res["rot"], args.steps)
    if "table" in res:
        out["fp32_tabulated_potential"] = brief(res["table"], args.steps)
    for name, v in res.get("variants", {}).items():
        out.setdefault("fp32_variants", {})[name] = brief(v, args.steps)
    traffic = os.path.join(ROOT, "profiles", f"traffic_m{n}.json")
    if os.path.exists(traffic):
        with open(traffic) as fh:
            out["roofline"]["traffic"] = json.load(fh).get("bytes_per_launch_fp32")
            out["roofline"]["traffic_unit"] = "bytes per launch (FP32 headline kernel)"
    if not args.no_warm_start:
        print("[bench] ground-state leg (network start vs TF start)", file=sys.stderr,
              flush=True)
        try:
            out["warm_start"] = warm_start_leg(args)
        except Exception as e:  # reported, never silently replaced
            out["warm_start"] = {"value": None, "error": str(e)[-500:]}
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        try:
            out["cpu_baseline"] = cpu_baseline(n)
        except Exception as e:  # reported, never silently replaced
            out["cpu_baseline"] = {"value": None, "error": str(e)[-500:]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
