"""Learned warm start for imaginary-time ground states.

A small MLP maps trap, interaction, rotation and atom number (w_perp, w_z, g, Omega,
log N) to a coarse 32^3 ground-state field (sqrt-density, cos phase, sin phase) on the
box [-12, 12)^3.  It is trained on ground states the batched imaginary-time solver itself
produces (Sobol-sampled parameters), on the GPU by libtbec's BF16-MFMA kernels.  Its
prediction is trilinearly upsampled to the fine grid and refined by FP32 imaginary time;
the plain Thomas-Fermi start is both the fallback and the judge.

continuation_solve relaxes a start on coarser grids of the same box first (larger stable
dtau, cheaper steps) and hands each level's state to the next by tbec_prolong: the slow
modes of a rotating ground state (a vortex lattice settling) take the same imaginary time
on every grid, so they are cheapest to relax on a coarse one.

    python -m trapbec_amd.warmstart [--samples 2000] [--fine 384] [--out DIR]
    python -m trapbec_amd.warmstart --phase continuation
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np

os.environ.setdefault("TBEC_PROGRESS", "20")

from trapbec_amd import autotune, initial, tbec  # noqa: E402

BOX = 24.0
RANGES = np.array([[0.8, 1.2],      # w_perp
                   [0.3, 1.5],      # w_z
                   [200.0, 4000.0], # g
                   [0.0, 0.8],      # Omega
                   [math.log(0.5), math.log(2.0)]])  # log N
COARSE = 32
# (tile, segment, XCD order) of the 384^3 rotating imaginary-time refinement: the fastest
# of autotune.tune(problem=...) for C5's problem on MI355X (278.9 us per stage; 0.829 ms
# per imaginary-time step against 0.902 with (2, 32, 1))
C5_TUNING = (2, 128, 2)


def sobol_params(n: int, seed: int) -> np.ndarray:
    """n scrambled-Sobol points over RANGES (rows: w_perp, w_z, g, Omega, log N)."""
    from scipy.stats import qmc
    u = qmc.Sobol(d=5, scramble=True, seed=seed).random(n)
    return RANGES[:, 0] + u * (RANGES[:, 1] - RANGES[:, 0])


def features(params: np.ndarray) -> np.ndarray:
    """Inputs scaled to [-1, 1]."""
    return (2.0 * (params - RANGES[:, 0]) / (RANGES[:, 1] - RANGES[:, 0]) - 1.0).astype(
        np.float32)


def problem(p, n, dtau):
    wp, wz, g, Om, logN = p
    shape = (n, n, n)
    return initial.Problem(shape, (BOX, BOX, BOX), (wp, wp, wz), g, Om, -dtau * 1j, False,
                           {"norm": math.exp(logN)})


def tf_start(pr, norm, seed=0, phase_amp=0.1):
    psi, _ = initial.tf_profile(pr.shape, pr.box, pr.omega, pr.g, norm=norm)
    return psi * initial.phase_noise(pr.shape, seed, phase_amp)


def coarse_field(psi: np.ndarray, cn: int = COARSE) -> np.ndarray:
    """(3, cn, cn, cn): block-averaged psi -> sqrt-density, cos phase, sin phase."""
    n = psi.shape[0]
    f = n // cn
    blk = psi.reshape(cn, f, cn, f, cn, f).mean(axis=(1, 3, 5))
    a = np.abs(blk)
    ph = np.angle(blk)
    return np.stack([a, np.cos(ph), np.sin(ph)]).astype(np.float32)


def generate(n_samples, seed=11, n=64, dtau=5e-3, tol=1e-8, max_steps=8000, chunk=250,
             log=print):
    """Ground states for Sobol parameters via the batched FP32 imaginary-time solver.
    Every instance starts from the TF profile with the same seeded phase noise, so the
    state reached (vortex positions included) is a function of the parameters alone."""
    params = sobol_params(n_samples, seed)
    X = features(params)
    Y = np.zeros((n_samples, 3 * COARSE ** 3), np.float32)
    stats = []
    for c0 in range(0, n_samples, chunk):
        idx = range(c0, min(n_samples, c0 + chunk))
        # the ground state at atom number N and coupling g is sqrt(N) times the one at
        # N = 1 and coupling g*N: every instance is solved at norm 1
        probs = [fine_problem(params[i], n, dtau) for i in idx]
        norms = [math.exp(params[i][4]) for i in idx]
        t0 = time.time()
        with tbec.Batch(probs, prec=32) as b:
            for k, i in enumerate(idx):
                b.upload(k, tf_start(probs[k], 1.0, seed=0))
            res = b.imag_solve(max_steps, tol, 1.0, 100)
            for k, i in enumerate(idx):
                Y[i] = coarse_field(b.download(k) * math.sqrt(norms[k])).ravel()
        stats += res
        log(f"[warmstart] samples {c0}..{idx[-1]}: {time.time() - t0:.1f}s, "
            f"converged {sum(r['converged'] for r in res)}/{len(res)}, "
            f"mean steps {np.mean([r['steps'] for r in res]):.0f}")
    return params, X, Y, stats


def fine_problem(p, n, dtau):
    """The refinement problem in units where the state is normalised to 1: g -> g N."""
    wp, wz, g, Om, logN = p
    return problem(np.array([wp, wz, g * math.exp(logN), Om, 0.0]), n, dtau)


def refine(pr, start, tol, max_steps, check_every=200, coarse=None, prec=32, tuning=None):
    """FP32 imaginary time to tol from a psi start (or a coarse field); returns stats.
    tuning: (tile, seg, remap) from autotune.tune(problem=pr)."""
    with tbec.Condensate.from_problem(pr, prec=prec) as c:
        if tuning is not None:
            c.set_tuning(tuning[0], tuning[1])
            c.set_xcd_remap(tuning[2])
        if coarse is not None:
            tbec.upload_coarse(c, coarse)
        else:
            c.upload(start)
        c.sync()
        t0 = time.time()
        r = c.imag_solve(max_steps, tol, 1.0, check_every)
        c.sync()
        r["wall_s"] = time.time() - t0
        d = c.diagnose()
        r["final_energy"], r["final_residual"] = d["energy"], d["residual"]
        return r


def refine_until(pr, e_target, r_target, max_steps, chunk=500, coarse=None, start=None,
                 prec=32, tuning=None):
    """Renormalised imaginary-time RK3 from a start until the energy and the GP residual
    are both at or below the targets (checked every `chunk` steps); returns stats."""
    with tbec.Condensate.from_problem(pr, prec=prec) as c:
        if tuning is not None:
            c.set_tuning(tuning[0], tuning[1])
            c.set_xcd_remap(tuning[2])
        if coarse is not None:
            tbec.upload_coarse(c, coarse)
        else:
            c.upload(start)
        c.sync()
        t0 = time.time()
        steps, d = 0, c.diagnose()
        last = t0
        while steps < max_steps and not (d["energy"] <= e_target and d["residual"] <= r_target):
            c.step(chunk, tbec.IMAG_RK3)
            steps += chunk
            d = c.diagnose()
            if time.time() - last > 20:
                print(f"[warmstart] {steps} steps: E {d['energy']:.10f} residual "
                      f"{d['residual']:.3e}", flush=True)
                last = time.time()
        return {"steps": steps, "reached": bool(d["energy"] <= e_target and
                                                 d["residual"] <= r_target),
                "energy": d["energy"], "residual": d["residual"],
                "wall_s": time.time() - t0}


def history(pr, max_steps, chunk=1000, coarse=None, start=None, prec=32, tuning=None):
    """Plain renormalised imaginary time for max_steps (no stop rule) from a start,
    recording (steps, E, GP residual) every `chunk` steps."""
    out = []
    with tbec.Condensate.from_problem(pr, prec=prec) as c:
        if tuning is not None:
            c.set_tuning(tuning[0], tuning[1])
            c.set_xcd_remap(tuning[2])
        if coarse is not None:
            tbec.upload_coarse(c, coarse)
        else:
            c.upload(start)
        d = c.diagnose()
        out.append((0, d["energy"], d["residual"]))
        t0 = last = time.time()
        steps = 0
        while steps < max_steps:
            c.step(chunk, tbec.IMAG_RK3)
            steps += chunk
            d = c.diagnose()
            out.append((steps, d["energy"], d["residual"]))
            if time.time() - last > 20:
                print(f"[warmstart] history {steps} steps: E {d['energy']:.10f} residual "
                      f"{d['residual']:.3e}", flush=True)
                last = time.time()
        wall = time.time() - t0
    return {"steps": steps, "wall_s": wall, "history": out}


def _start_handle(pr, coarse=None, start=None, prec=32, tuning=None):
    c = tbec.Condensate.from_problem(pr, prec=prec)
    if tuning is not None:
        c.set_tuning(tuning[0], tuning[1])
        c.set_xcd_remap(tuning[2])
    if coarse is not None:
        tbec.upload_coarse(c, coarse)
    else:
        c.upload(start)
    return c


def run_to_state(c, max_steps, chunk, e_target=None, r_target=None, log=print, tag=""):
    """Renormalised imaginary-time steps on handle c in chunks of `chunk`, with the FP64
    diagnostics (energy, GP residual) after every chunk, until both are at or below the
    targets (when given) or max_steps.  Returns steps, wall time (first step to last
    diagnostics, device-synchronised) and the (steps, E, residual) record."""
    c.sync()
    t0 = last = time.time()
    d = c.diagnose()
    rec = [(0, d["energy"], d["residual"])]
    steps = 0

    def reached(dd):
        return e_target is not None and dd["energy"] <= e_target and dd["residual"] <= r_target

    while steps < max_steps and not reached(d):
        n = min(chunk, max_steps - steps)
        c.step(n, tbec.IMAG_RK3)
        steps += n
        d = c.diagnose()
        rec.append((steps, d["energy"], d["residual"]))
        if time.time() - last > 20:
            log(f"[warmstart] {tag} {steps} steps: E {d['energy']:.10f} residual "
                f"{d['residual']:.3e}")
            last = time.time()
    return {"steps": steps, "wall_s": time.time() - t0, "reached": reached(d),
            "energy": d["energy"], "residual": d["residual"], "record": rec}


def equal_state_comparison(pr, coarse, warm_steps, cap_factor=5.0, chunk=1000,
                           tuning=None, infer_s=0.0, log=print):
    """The network start runs `warm_steps` steps and fixes the target: its energy and GP
    residual.  The TF start (the same seeded TF profile and phase noise the training
    states start from) then runs until it reaches both targets or cap_factor times the
    network start's steps.  Both runs use the same step loop and diagnostics cadence, so
    steps and wall times compare directly.  Returns the two runs and the speed-up (a
    lower bound when the TF start does not reach the target within the cap)."""
    t0 = time.time()
    c = _start_handle(pr, coarse=coarse, tuning=tuning)
    up_s = time.time() - t0
    warm = run_to_state(c, warm_steps, chunk, log=log, tag="network start")
    c.close()
    warm["wall_s"] += up_s + infer_s
    cap = int(cap_factor * warm["steps"])
    t0 = time.time()
    c = _start_handle(pr, start=tf_start(pr, 1.0, seed=0), tuning=tuning)
    up_s = time.time() - t0
    tf = run_to_state(c, cap, chunk, e_target=warm["energy"], r_target=warm["residual"],
                      log=log, tag="TF start")
    c.close()
    tf["wall_s"] += up_s
    out = {"network_start": {k: v for k, v in warm.items() if k != "record"},
           "tf_start": {k: v for k, v in tf.items() if k != "record"},
           "tf_cap_steps": cap,
           "target": {"energy": warm["energy"], "residual": warm["residual"]}}
    if tf["reached"]:
        out["speedup_steps"] = tf["steps"] / warm["steps"]
        out["speedup_wall"] = tf["wall_s"] / warm["wall_s"]
    else:
        out["speedup_steps_lower_bound"] = cap / warm["steps"]
        out["speedup_wall_lower_bound"] = tf["wall_s"] / warm["wall_s"]
    out["records"] = {"network_start": warm["record"], "tf_start": tf["record"]}
    return out


def train_network(samples=2000, train_n=64, gen_max=8000, gen_tol=1e-8, epochs=300,
                  batch=64, lr=1e-3, seed=11, log=print):
    """Training data from the batched solver, then the network trained on the GPU.
    Returns (mlp handle, output scale, stats); the caller closes the handle."""
    t0 = time.time()
    params, X, Y, stats = generate(samples, seed=seed, n=train_n, tol=gen_tol,
                                   max_steps=gen_max, log=log)
    t_gen = time.time() - t0
    scale = np.ones(Y.shape[1], np.float32)
    scale[:COARSE ** 3] = 1.0 / Y[:, :COARSE ** 3].max()     # sqrt-density to O(1)
    m = tbec.MLP(5, 512, 4, 3 * COARSE ** 3, seed=1)
    n_tr = int(0.9 * samples)
    m.set_data(X[:n_tr], Y[:n_tr] * scale)
    t1 = time.time()
    hist = m.train(epochs, batch, lr, seed=2)
    t_train = time.time() - t1
    val = m.forward(X[n_tr:]) / scale
    info = {"samples": samples, "train_grid": train_n, "gen_s": t_gen, "train_s": t_train,
            "gen_converged": int(sum(s["converged"] for s in stats)),
            "epochs": epochs, "loss_first_last": [float(hist[0]), float(hist[-1])],
            "val_mse": float(np.mean((val - Y[n_tr:]) ** 2)),
            "val_mse_mean_predictor": float(np.mean((Y[:n_tr].mean(axis=0) - Y[n_tr:]) ** 2))}
    log(f"[warmstart] data {t_gen:.1f} s ({info['gen_converged']}/{samples} converged), "
        f"training {t_train:.1f} s, val MSE {info['val_mse']:.4f} (mean predictor "
        f"{info['val_mse_mean_predictor']:.4f})")
    return m, scale, info


def predict(m, scale, test):
    """Coarse field (3, 32, 32, 32) for parameters `test`, at norm 1, and the time taken."""
    t0 = time.time()
    coarse = (m.forward(features(np.asarray(test)[None]))[0] / scale).reshape(
        3, COARSE, COARSE, COARSE)
    coarse[0] /= math.sqrt(math.exp(test[4]))   # the refinement runs at norm 1
    return coarse, time.time() - t0


def first_reaching(hist, e_target, r_target=None):
    """First recorded step count with E <= e_target (and residual <= r_target)."""
    for st, e, r in hist:
        if e <= e_target and (r_target is None or r <= r_target):
            return st
    return None


RK3_REAL_LIMIT = 2.5127   # |z| where 1 + z + z^2/2 + z^3/6 leaves the unit disc, z < 0


def max_rate(pr, shape=None) -> float:
    """Upper bound of the imaginary-time operator's largest eigenvalue on `shape` (default
    pr.shape): the 6th-order -1/2 Laplacian at the grid's highest wavenumber
    (1088/360 / h^2 per axis: the stencil's symbol at theta = pi), the trap at the box
    corner, |Omega| times the box's largest radius times the 6th-order first derivative's
    bound (110/60 / h, its summed |coefficients|), and 2 mu_TF for the nonlinearity
    (mu_TF of the norm-1 Thomas-Fermi profile)."""
    shape = shape or pr.shape
    h = [L / n for L, n in zip(pr.box, shape)]
    kin = sum(1088.0 / 360.0 / hh ** 2 for hh in h)
    v = 0.5 * sum((w * L / 2) ** 2 for w, L in zip(pr.omega, pr.box))
    rot = abs(pr.Omega) * math.hypot(pr.box[0] / 2, pr.box[1] / 2) * (110.0 / 60.0) / \
        min(h[0], h[1])
    wbar3 = pr.omega[0] * pr.omega[1] * pr.omega[2]
    mu = 0.5 * (15.0 * pr.g * wbar3 / (4.0 * math.pi)) ** 0.4
    return kin + v + rot + 2.0 * mu


def continuation_levels(pr, factors=(3, 2), safety=0.95):
    """Coarser grids of pr's box for a coarse-to-fine ground-state solve, coarsest first:
    shape // factor per level, each with the largest dtau the stability bound allows
    (safety * 2.5127 / max_rate), never below pr's own dtau."""
    dtau = abs(complex(pr.dt))
    out = []
    for f in factors:
        shape = tuple(max(2, n // f) for n in pr.shape)
        out.append((shape, max(dtau, safety * RK3_REAL_LIMIT / max_rate(pr, shape))))
    return out


def continuation_solve(pr, coarse=None, factors=(3, 2), tol=1e-10, gate=5e-4,
                       coarse_gate=1e-4, max_steps=400000, check_every=1000, tuning=None,
                       safety=0.95, keep=False, log=print):
    """Imaginary-time ground state of pr on a ladder of grids: the coarsest level starts
    from the network's coarse field (`coarse`, trilinearly upsampled) or, when None, from
    the TF profile with the seeded phase noise of tf_start; each level runs the gated stop
    rule (relative energy change below tol for 8 steps, GP residual at or below
    coarse_gate; at pr's own grid and dtau, at or below gate) and hands its state to the
    next by tbec_prolong.  Returns per-level steps, wall times, energies and residuals and
    the total wall time; keep=True also returns the final handle (the caller closes it)."""
    levels = continuation_levels(pr, factors, safety) + [(tuple(pr.shape), abs(complex(pr.dt)))]
    t0 = time.time()
    prev, recs = None, []
    for k, (shape, dtau) in enumerate(levels):
        fine = k == len(levels) - 1
        p = initial.Problem(shape, pr.box, pr.omega, pr.g, pr.Omega, -dtau * 1j, False,
                            dict(pr.meta))
        t1 = time.time()
        c = tbec.Condensate.from_problem(p, prec=32)
        if fine and tuning is not None:
            c.set_tuning(tuning[0], tuning[1])
            c.set_xcd_remap(tuning[2])
        if prev is not None:
            tbec.prolong(c, prev)
            prev.close()
        elif coarse is not None:
            tbec.upload_coarse(c, coarse)
        else:
            c.upload(tf_start(p, 1.0, seed=0))
        c.set_residual_gate(gate if fine else coarse_gate)
        r = c.imag_solve(max_steps, tol, 1.0, check_every)
        d = c.diagnose()
        recs.append({"shape": list(shape), "dtau": dtau, "steps": int(r["steps"]),
                     "converged": bool(r["converged"]), "wall_s": time.time() - t1,
                     "energy": d["energy"], "residual": d["residual"]})
        log(f"[continuation] {'x'.join(map(str, shape))} dtau {dtau:.3g}: "
            f"{int(r['steps'])} steps {recs[-1]['wall_s']:.1f} s E {d['energy']:.10f} "
            f"residual {d['residual']:.3e}{'' if r['converged'] else ' (step cap)'}")
        prev = c
    out = {"levels": recs, "wall_s": time.time() - t0, "energy": recs[-1]["energy"],
           "residual": recs[-1]["residual"], "converged": recs[-1]["converged"],
           "fine_steps": recs[-1]["steps"]}
    if keep:
        return out, prev
    prev.close()
    return out


def converge_network_start(pr, coarse, tol, gate, max_steps, check_every=1000, tuning=None,
                           infer_s=0.0):
    """The network start refined to the stop rule: relative energy change below tol for 8
    consecutive steps with the GP residual at or below `gate` (tbec_imag_solve with the
    residual gate), at most max_steps.  Returns steps, wall time (network inference,
    upsampling and the solve), energy and residual of the final state."""
    t0 = time.time()
    c = _start_handle(pr, coarse=coarse, tuning=tuning)
    c.set_residual_gate(gate)
    c.sync()
    r = c.imag_solve(max_steps, tol, 1.0, check_every)
    c.sync()
    wall = time.time() - t0 + infer_s
    d = c.diagnose()
    c.close()
    return {"steps": int(r["steps"]), "converged": bool(r["converged"]), "wall_s": wall,
            "energy": d["energy"], "residual": d["residual"], "tol": tol, "gate": gate}


def converge_main(a):
    """Both starts to the converged state at C5's size, in two phases that fit one GPU
    call each: `--phase network` trains the network and runs its start to the gated stop
    rule; `--phase tf --target E R` runs the TF start until its energy and GP residual are
    at or below E and R (the network start's final state), or --max-steps.
    `--phase continuation` runs the coarse-to-fine solve (continuation_solve) from the
    network's field and from the TF start."""
    test = np.array(a.test)
    pr = fine_problem(test, a.fine, a.fine_dtau)
    tuning = C5_TUNING
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.phase == "continuation":
        # the coarse-to-fine solve from the network's field and from the TF start
        m, scale, info = train_network(samples=a.samples, train_n=a.train_n, epochs=a.epochs,
                                       log=log)
        try:
            coarse, infer_s = predict(m, scale, test)
        finally:
            m.close()
        net = continuation_solve(pr, coarse=coarse, tol=a.tol, gate=a.gate,
                                 max_steps=a.max_steps, tuning=tuning, log=log)
        net["wall_s"] += infer_s
        tf = continuation_solve(pr, tol=a.tol, gate=a.gate, max_steps=a.max_steps,
                                tuning=tuning, log=log)
        out = {"phase": "continuation", "network": info, "inference_s": infer_s,
               "network_start": net, "tf_start": tf}
    elif a.phase == "network":
        m, scale, info = train_network(samples=a.samples, train_n=a.train_n, epochs=a.epochs,
                                       log=log)
        try:
            coarse, infer_s = predict(m, scale, test)
        finally:
            m.close()
        os.environ.setdefault("TBEC_PROGRESS", "20")
        r = converge_network_start(pr, coarse, a.tol, a.gate, a.max_steps, tuning=tuning,
                                   infer_s=infer_s)
        out = {"phase": "network", "network": info, "network_start": r}
    else:
        c = _start_handle(pr, start=tf_start(pr, 1.0, seed=0), tuning=tuning)
        r = run_to_state(c, a.max_steps, 1000, e_target=a.target[0], r_target=a.target[1],
                         log=log, tag="TF start")
        c.close()
        rec = r.pop("record")
        out = {"phase": "tf", "target": {"energy": a.target[0], "residual": a.target[1]},
               "tf_start": r, "record_every_10k": [x for x in rec if x[0] % 10000 == 0]}
    out.update({"test_params": test.tolist(), "grid": a.fine, "dtau": a.fine_dtau})
    tag = "_".join(f"{v:g}" for v in a.test)
    os.makedirs(a.out, exist_ok=True)
    with open(os.path.join(a.out, f"converge_{a.phase}_{tag}.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--train-n", type=int, default=64)
    ap.add_argument("--fine", type=int, default=384)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--fine-dtau", type=float, default=8e-4)
    ap.add_argument("--max-steps", type=int, default=200000)
    ap.add_argument("--test", type=float, nargs=5, default=[1.0, 0.8, 3000.0, 0.5, 0.0],
                    help="held-out parameters: w_perp w_z g Omega log_N")
    ap.add_argument("--history", type=int, default=0,
                    help="also run both starts this many steps without the stop rule and "
                         "compare steps to equal energy along the way")
    ap.add_argument("--out", default="gpurun_out/warmstart")
    ap.add_argument("--phase", choices=["network", "tf", "continuation"], default=None,
                    help="converged comparison in two phases, or the coarse-to-fine solve "
                         "from both starts (see converge_main)")
    ap.add_argument("--gate", type=float, default=5e-4,
                    help="GP-residual gate of the stop rule for --phase network")
    ap.add_argument("--target", type=float, nargs=2, default=None,
                    help="--phase tf: energy and GP residual to reach")
    a = ap.parse_args()
    if a.phase:
        converge_main(a)
        return
    os.makedirs(a.out, exist_ok=True)
    t0 = time.time()
    params, X, Y, stats = generate(a.samples, n=a.train_n,
                                   log=lambda m: print(m, flush=True))
    t_gen = time.time() - t0
    scale = np.ones(Y.shape[1], np.float32)
    amp = Y[:, :COARSE ** 3].max()
    scale[:COARSE ** 3] = 1.0 / amp          # sqrt-density channel to O(1)
    with tbec.MLP(5, 512, 4, 3 * COARSE ** 3, seed=1) as m:
        n_tr = int(0.9 * a.samples)
        m.set_data(X[:n_tr], Y[:n_tr] * scale)
        t1 = time.time()
        hist = m.train(a.epochs, a.batch, a.lr, seed=2)
        t_train = time.time() - t1
        print(f"[warmstart] trained {a.epochs} epochs in {t_train:.1f} s, loss "
              f"{hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
        val_pred = m.forward(X[n_tr:]) / scale
        val_mse = float(np.mean((val_pred - Y[n_tr:]) ** 2))
        base_mse = float(np.mean((Y[:n_tr].mean(axis=0) - Y[n_tr:]) ** 2))
        test = np.array(a.test)
        coarse = (m.forward(features(test[None]))[0] / scale).reshape(3, COARSE, COARSE, COARSE)
    # the network predicts the state at atom number N; the refinement runs at norm 1
    coarse[0] /= math.sqrt(math.exp(test[4]))
    pr = fine_problem(test, a.fine, a.fine_dtau)
    e = autotune.tune(None, 32, problem=pr)
    tuning = (e["tile"], e["seg"], e["remap"])
    print(f"[warmstart] tuning {tuning}: {e['us_per_stage']:.1f} us/stage", flush=True)
    print(f"[warmstart] refine {a.fine}^3 from the TF start", flush=True)
    r_tf = refine(pr, tf_start(pr, 1.0, seed=0), a.tol, a.max_steps, tuning=tuning)
    print(f"[warmstart] TF start: {r_tf['steps']} steps {r_tf['wall_s']:.1f} s", flush=True)
    # equal final energy and GP residual: steps the warm start needs to reach the state
    # the TF start ended in (converged or at its step limit)
    e_t, r_t = r_tf["final_energy"], r_tf["final_residual"]
    r_ml = refine_until(pr, e_t, r_t, a.max_steps, coarse=coarse, tuning=tuning)
    print(f"[warmstart] MLP start reaches E <= {e_t:.10f}, residual <= {r_t:.3e}: "
          f"{r_ml['steps']} steps {r_ml['wall_s']:.1f} s", flush=True)
    r_ml_tol = None
    if r_tf["converged"]:
        # both to the stop rule: the north-star comparison when the TF start gets there
        r_ml_tol = refine(pr, None, a.tol, a.max_steps, coarse=coarse, tuning=tuning)
        print(f"[warmstart] MLP start to tol {a.tol:g}: {r_ml_tol['steps']} steps "
              f"{r_ml_tol['wall_s']:.1f} s, E {r_ml_tol['final_energy']:.10f}", flush=True)
    out = {"samples": a.samples, "gen_s": t_gen, "train_s": t_train,
           "train_loss_first_last": [float(hist[0]), float(hist[-1])],
           "val_mse": val_mse, "val_mse_mean_predictor": base_mse,
           "gen_converged": int(sum(s["converged"] for s in stats)),
           "fine_grid": a.fine, "test_params": test.tolist(), "tf_start": r_tf,
           "mlp_start": r_ml,
           "criterion": "steps until E and GP residual reach the TF start's final values",
           "speedup_steps": r_tf["steps"] / max(1, r_ml["steps"]),
           "speedup_wall": r_tf["wall_s"] / max(1e-9, r_ml["wall_s"]),
           "mlp_start_to_tol": r_ml_tol,
           "speedup_steps_to_tol": (r_tf["steps"] / max(1, r_ml_tol["steps"])
                                    if r_ml_tol and r_ml_tol["converged"] else None)}
    if a.history:
        h_tf = history(pr, a.history, start=tf_start(pr, 1.0, seed=0), tuning=tuning)
        h_ml = history(pr, a.history, coarse=coarse, tuning=tuning)
        e_best = min(e for _, e, _ in h_tf["history"] + h_ml["history"])
        ladder = {}
        for rel in (1e-2, 1e-3, 1e-4, 1e-5):
            et = e_best * (1 + rel)
            ladder[f"{rel:g}"] = {"E_target": et,
                                  "tf_steps": first_reaching(h_tf["history"], et),
                                  "mlp_steps": first_reaching(h_ml["history"], et)}
            print(f"[warmstart] E <= E_min*(1+{rel:g}): TF {ladder[f'{rel:g}']['tf_steps']} "
                  f"MLP {ladder[f'{rel:g}']['mlp_steps']} steps", flush=True)
        out["history"] = {"E_min": e_best, "steps_to_energy": ladder,
                          "tf": h_tf, "mlp": h_ml}
    tag = "_".join(f"{v:g}" for v in a.test)
    with open(os.path.join(a.out, f"warmstart_{tag}.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
