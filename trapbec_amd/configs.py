"""The workload configurations, run end to end on the GPU, with their reported figures.

    python -m trapbec_amd.configs c2|c3|c4 [--out DIR] [--n N]

c2: real time at 256^3 in FP64 and FP32 from the same state: psi relative L2 error,
    norm / energy drift and vortex-core displacement of FP32 against FP64.
c3: imaginary-time ground state in a rotating trap at 256^3, FP32 bulk with FP64
    reductions against the FP64 run: energy error, vortex-lattice count.
c4: 48 condensates of uneven grids to energy convergence: one batched solve against
    solving them one at a time on the same GPU; --continuation solves each on its grid
    halved first (coarse-to-fine).
"""
from __future__ import annotations

import argparse
import json
import os
import time

os.environ.setdefault("TBEC_PROGRESS", "20")

import numpy as np

from trapbec_amd import initial, tbec


def _cores(c, dens_min, orientation=0):
    return c.vortex_cores(dens_min, orientation)


def _match_displacement(a, b, h):
    """Mean / max distance (in units of x) from each core of a to the nearest core of b
    with the same charge and z index; plaquette positions in grid units times h."""
    if len(a) == 0 or len(b) == 0:
        return None, None
    d = []
    for i, j, k, q in a:
        sel = b[(b[:, 2] == k) & (b[:, 3] == q)]
        if len(sel) == 0:
            continue
        dd = np.sqrt((sel[:, 0] - i) ** 2 + (sel[:, 1] - j) ** 2).min() * h
        d.append(dd)
    if not d:
        return None, None
    return float(np.mean(d)), float(np.max(d))


def run_c2(n=256, steps=2000):
    pr, psi0 = initial.config_c2(n=n)
    out = {"config": "c2", "grid": n, "steps": steps}
    finals, diags = {}, {}
    for prec in (64, 32):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.upload(psi0)
            d0 = c.diagnose()
            t0 = time.time()
            c.step(steps)
            c.sync()
            wall = time.time() - t0
            d1 = c.diagnose()
            finals[prec] = c.download().astype(np.complex128)
            cores = _cores(c, 1e-5 * float(np.max(np.abs(psi0)) ** 2))
        diags[prec] = {"norm0": d0["norm"], "norm1": d1["norm"], "E0": d0["energy"],
                       "E1": d1["energy"], "wall_s": wall, "n_cores_z": int(len(cores))}
        finals[(prec, "cores")] = cores
    ref = finals[64]
    out["psi_rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(finals[32] - ref) / np.linalg.norm(ref))
    out["psi_rel_max_fp32_vs_fp64"] = float(np.abs(finals[32] - ref).max() / np.abs(ref).max())
    for prec in (64, 32):
        d = diags[prec]
        d["norm_drift"] = abs(d["norm1"] - d["norm0"]) / d["norm0"]
        d["energy_drift"] = abs(d["E1"] - d["E0"]) / abs(d["E0"])
        out[f"fp{prec}"] = d
    h = pr.box[0] / n
    mean_d, max_d = _match_displacement(finals[(32, "cores")], finals[(64, "cores")], h)
    out["vortex_core_displacement_mean"] = mean_d
    out["vortex_core_displacement_max"] = max_d
    out["grid_spacing"] = h
    return out


def run_c3(n=256, tol=1e-10, max_steps=200000, check_every=200, gate=0.0):
    pr, psi0 = initial.config_c3(n=n)
    out = {"config": "c3", "grid": n, "tol": tol, "residual_gate": gate}
    res = {}
    from trapbec_amd import autotune
    for prec in (64, 32):
        e = autotune.tune(None, prec, problem=pr)
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.set_tuning(e["tile"], e["seg"])
            c.set_xcd_remap(e["remap"])
            c.set_residual_gate(gate)
            c.upload(psi0)
            print(f"[c3] fp{prec} imaginary-time solve {pr.shape}, tile {e['tile']} seg "
                  f"{e['seg']} remap {e['remap']}: {e['us_per_stage']:.1f} us/stage",
                  flush=True)
            t0 = time.time()
            r = c.imag_solve(max_steps, tol, 1.0, check_every)
            c.sync()
            r["wall_s"] = time.time() - t0
            d = c.diagnose()
            r["final_energy"] = d["energy"]
            r["final_residual"] = d["residual"]
            psi = c.download()
            dens = float(np.max(np.abs(psi)) ** 2)
            cores = c.vortex_cores(1e-5 * dens, 0)
            mid = n // 2
            r["vortex_count_midplane"] = int(np.sum(cores[:, 2] == mid)) if len(cores) else 0
            r["vortex_charge_sum_midplane"] = int(np.sum(cores[cores[:, 2] == mid][:, 3])) \
                if len(cores) else 0
        res[prec] = r
    out["fp64"], out["fp32"] = res[64], res[32]
    out["energy_rel_error_fp32_vs_fp64"] = abs(res[32]["final_energy"] - res[64]["final_energy"]) \
        / abs(res[64]["final_energy"])
    return out


def run_c4(n_inst=48, tol=1e-10, max_steps=20000, check_every=100, prec=32, gate=0.0):
    """gate > 0: the stop rule also requires the GP residual at or below it (a rotating
    instance cannot stop in its vortex-free saddle)."""
    from trapbec_amd import autotune
    probs = initial.config_c4(n_inst=n_inst)
    out = {"config": "c4", "instances": n_inst, "tol": tol, "max_steps": max_steps,
           "residual_gate": gate,
           "total_points": int(sum(np.prod(p.shape) for p, _ in probs))}
    # tile shapes from the autotuner: one (tile, seg) for the ensemble, one per grid
    # shape for the one-at-a-time solves
    tb = autotune.tune_batch([p for p, _ in probs], prec=prec)
    print(f"[c4] batch tuning: tile {tb['tile']} seg {tb['seg']} "
          f"{tb['ms_per_step']:.3f} ms/step", flush=True)
    single = {}
    for shape in sorted({tuple(p.shape) for p, _ in probs}):
        first = next(p for p, _ in probs if tuple(p.shape) == shape)
        e = autotune.tune(None, prec, problem=first)
        single[shape] = (e["tile"], e["seg"], e["remap"])
        print(f"[c4] single tuning {shape}: tile {e['tile']} seg {e['seg']} "
              f"remap {e['remap']}", flush=True)
    out["tuning"] = {"batch": [tb["tile"], tb["seg"]],
                     "single": {"x".join(map(str, k)): v for k, v in single.items()}}
    t0 = time.time()
    with tbec.Batch([p for p, _ in probs], prec=prec) as b:
        b.set_tuning(tb["tile"], tb["seg"])
        b.set_residual_gate(gate)
        for i, (_, psi0) in enumerate(probs):
            b.upload(i, psi0)
        t1 = time.time()
        res_b = b.imag_solve(max_steps, tol, 1.0, check_every)
        t_batch = time.time() - t1
        fin = [b.download(i) for i in range(len(probs))]
    out["batched"] = {"wall_s": t_batch, "setup_s": t1 - t0,
                      "converged": int(sum(r["converged"] for r in res_b)),
                      "steps": [int(r["steps"]) for r in res_b],
                      "energy": [r["energy"] for r in res_b],
                      "residual": [r["residual"] for r in res_b]}
    # vortex cores in the midplane of each final state (z-normal plaquettes)
    cores = []
    for (pr, _), psi in zip(probs, fin):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.upload(psi)
            v = c.vortex_cores(1e-5 * float(np.max(np.abs(psi)) ** 2), 0)
            mid = pr.shape[2] // 2
            cores.append(int(np.sum(v[:, 2] == mid)) if len(v) else 0)
    out["batched"]["midplane_cores"] = cores
    del fin
    print(f"[c4] batched: {t_batch:.1f} s, {out['batched']['converged']}/{n_inst} converged",
          flush=True)
    seq = []
    t_seq = 0.0
    for k, (pr, psi0) in enumerate(probs):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            tile, seg, remap = single[tuple(pr.shape)]
            c.set_tuning(tile, seg)
            c.set_xcd_remap(remap)
            c.set_residual_gate(gate)
            c.upload(psi0)
            c.sync()
            t1 = time.time()
            r = c.imag_solve(max_steps, tol, 1.0, check_every)
            c.sync()
            t_seq += time.time() - t1
            seq.append(r)
        print(f"[c4] sequential {k + 1}/{n_inst} {pr.shape}: {int(r['steps'])} steps, "
              f"{t_seq:.1f} s so far", flush=True)
    out["sequential"] = {"wall_s": t_seq,
                         "converged": int(sum(r["converged"] for r in seq)),
                         "steps": [int(r["steps"]) for r in seq]}
    out["steps_identical"] = [int(a["steps"]) for a in res_b] == [int(a["steps"]) for a in seq]
    pt_steps = sum(int(np.prod(p.shape)) * int(r["steps"]) for (p, _), r in zip(probs, res_b))
    out["batched_gpt_steps_per_s"] = pt_steps / t_batch / 1e9
    out["sequential_gpt_steps_per_s"] = pt_steps / t_seq / 1e9
    out["speedup"] = t_seq / t_batch
    return out


def _midplane_cores(pr, psi, prec):
    with tbec.Condensate.from_problem(pr, prec=prec) as c:
        c.upload(psi)
        v = c.vortex_cores(1e-5 * float(np.max(np.abs(psi)) ** 2), 0)
    return int(np.sum(v[:, 2] == pr.shape[2] // 2)) if len(v) else 0


def run_c4_continuation(n_inst=48, tol=1e-10, gate=5e-4, coarse_gate=1e-4, factor=2,
                        max_steps=100000, coarse_max=400000, check_every=100, prec=32,
                        sequential=True, fine_bound=False, save=None):
    """C4 by the coarse-to-fine solve.  Each instance's start is interpolated onto its
    grid divided by `factor` (dtau there from the stability bound), solved to the gated
    stop rule with coarse_gate, interpolated back and solved to the gated rule with
    `gate` at its own grid, with its own dtau (1e-3) or, with fine_bound, the stability
    bound's there too.  Once as two batches (tbec_batch_prolong between them), then
    (sequential=True) once one instance at a time (tbec_prolong).  save(out), when given,
    is called with the results so far after the batched solve."""
    from trapbec_amd import autotune
    from trapbec_amd.warmstart import RK3_REAL_LIMIT, max_rate
    probs = initial.config_c4(n_inst=n_inst)
    if fine_bound:
        # the configuration fixes the grid (dx = 0.2), not dtau: take the largest step
        # the RK3 stability bound allows on that grid (the converged state does not
        # depend on it; the per-step stop rule is stricter per unit of imaginary time)
        probs = [(initial.Problem(p.shape, p.box, p.omega, p.g, p.Omega,
                                  -max(abs(complex(p.dt)),
                                       0.95 * RK3_REAL_LIMIT / max_rate(p)) * 1j,
                                  False, dict(p.meta)), psi0) for p, psi0 in probs]
    coarse = []
    for p, _ in probs:
        shape = tuple(max(2, n // factor) for n in p.shape)
        dt = max(abs(complex(p.dt)), 0.95 * RK3_REAL_LIMIT / max_rate(p, shape))
        coarse.append(initial.Problem(shape, p.box, p.omega, p.g, p.Omega, -dt * 1j, False,
                                      dict(p.meta)))
    out = {"config": "c4", "method": "coarse-to-fine", "instances": n_inst, "tol": tol,
           "gate": gate, "coarse_gate": coarse_gate, "factor": factor,
           "max_steps": max_steps, "coarse_max": coarse_max,
           "coarse_dtau": [abs(complex(c.dt)) for c in coarse],
           "fine_dtau": [abs(complex(p.dt)) for p, _ in probs],
           "total_points": int(sum(np.prod(p.shape) for p, _ in probs))}
    # each ensemble gets its own tuning: the coarse grids hold an eighth of the points and
    # prefer shorter segments (C4: tile (64,16,2) with segment 16 at 0.247 ms per step
    # against 0.307 with the fine ensemble's segment 32)
    tb = autotune.tune_batch([p for p, _ in probs], prec=prec)
    print(f"[c4] batch tuning: tile {tb['tile']} seg {tb['seg']}", flush=True)
    tc = autotune.tune_batch(coarse, prec=prec)
    print(f"[c4] coarse batch tuning: tile {tc['tile']} seg {tc['seg']}", flush=True)
    out["tuning"] = {"fine": [tb["tile"], tb["seg"]], "coarse": [tc["tile"], tc["seg"]]}
    t0 = time.time()
    with tbec.Batch([p for p, _ in probs], prec=prec) as bf, \
            tbec.Batch(coarse, prec=prec) as bc:
        bf.set_tuning(tb["tile"], tb["seg"])
        bc.set_tuning(tc["tile"], tc["seg"])
        # the last few unfinished instances run on shorter segments (C4: one 96^3 instance
        # runs the last ~70% of its 1.5 M fine steps alone)
        bf.set_tail_segment(True)
        bc.set_tail_segment(True)
        for i, (_, psi0) in enumerate(probs):
            bf.upload(i, psi0)
        t1 = time.time()
        tbec.batch_prolong(bc, bf)
        bc.set_residual_gate(coarse_gate)
        rc = bc.imag_solve(coarse_max, tol, 1.0, check_every)
        t2 = time.time()
        print(f"[c4] coarse batch: {t2 - t1:.1f} s, "
              f"{sum(r['converged'] for r in rc)}/{n_inst} converged", flush=True)
        tbec.batch_prolong(bf, bc)
        bf.set_residual_gate(gate)
        rf = bf.imag_solve(max_steps, tol, 1.0, check_every)
        t3 = time.time()
        fin = [bf.download(i) for i in range(n_inst)]
    out["batched"] = {"wall_s": t3 - t1, "coarse_s": t2 - t1, "fine_s": t3 - t2,
                      "setup_s": t1 - t0,
                      "converged": int(sum(r["converged"] for r in rf)),
                      "coarse_converged": int(sum(r["converged"] for r in rc)),
                      "coarse_steps": [int(r["steps"]) for r in rc],
                      "steps": [int(r["steps"]) for r in rf],
                      "energy": [r["energy"] for r in rf],
                      "residual": [r["residual"] for r in rf],
                      "midplane_cores": [_midplane_cores(p, psi, prec)
                                         for (p, _), psi in zip(probs, fin)]}
    del fin
    print(f"[c4] batched: {t3 - t1:.1f} s, {out['batched']['converged']}/{n_inst} converged",
          flush=True)
    if save:
        save(out)
    if not sequential:
        return out
    t_seq, seq = 0.0, []
    for k, ((p, psi0), pc) in enumerate(zip(probs, coarse)):
        with tbec.Condensate.from_problem(p, prec=prec) as cf, \
                tbec.Condensate.from_problem(pc, prec=prec) as cc:
            cf.upload(psi0)
            cf.sync()
            t1 = time.time()
            tbec.prolong(cc, cf)
            cc.set_residual_gate(coarse_gate)
            r1 = cc.imag_solve(coarse_max, tol, 1.0, check_every)
            tbec.prolong(cf, cc)
            cf.set_residual_gate(gate)
            r2 = cf.imag_solve(max_steps, tol, 1.0, check_every)
            cf.sync()
            t_seq += time.time() - t1
            seq.append((r1, r2))
        print(f"[c4] sequential {k + 1}/{n_inst} {p.shape}: {int(r1['steps'])} coarse + "
              f"{int(r2['steps'])} fine steps, {t_seq:.1f} s so far", flush=True)
    out["sequential"] = {"wall_s": t_seq,
                         "converged": int(sum(r2["converged"] for _, r2 in seq)),
                         "coarse_steps": [int(r1["steps"]) for r1, _ in seq],
                         "steps": [int(r2["steps"]) for _, r2 in seq],
                         "energy": [r2["energy"] for _, r2 in seq]}
    out["speedup"] = t_seq / out["batched"]["wall_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=["c2", "c3", "c4"])
    ap.add_argument("--out", default="gpurun_out/configs")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--gate", type=float, default=0.0,
                    help="c3, c4: GP-residual gate of the stop rule (0: none)")
    ap.add_argument("--continuation", action="store_true",
                    help="c4: the coarse-to-fine solve (grid / 2 first), gate 5e-4 unless "
                         "--gate is given")
    ap.add_argument("--coarse-max", type=int, default=400000,
                    help="c4 --continuation: step cap of the coarse level")
    ap.add_argument("--batch-only", action="store_true",
                    help="c4 --continuation: skip the one-at-a-time comparison")
    ap.add_argument("--fine-bound", action="store_true",
                    help="c4 --continuation: dtau of each instance's own grid from the RK3 "
                         "stability bound (default: 1e-3)")
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    name = f"{a.config}_continuation" if a.continuation else a.config

    def save(r):
        with open(os.path.join(a.out, f"{name}.json"), "w") as f:
            json.dump(r, f, indent=1)

    if a.config == "c2":
        r = run_c2(n=a.n)
    elif a.config == "c3":
        r = run_c3(n=a.n, max_steps=a.max_steps or 200000, gate=a.gate)
    elif a.continuation:
        r = run_c4_continuation(max_steps=a.max_steps or 100000, gate=a.gate or 5e-4,
                                coarse_max=a.coarse_max, sequential=not a.batch_only,
                                fine_bound=a.fine_bound, save=save)
    else:
        r = run_c4(max_steps=a.max_steps or 20000, gate=a.gate)
    save(r)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
