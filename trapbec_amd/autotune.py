"""Tile-shape autotuner for the fused stage kernel.

The stage kernel is compiled for five (TZ, TY, VZ) tile shapes per precision and takes
the x-segment length SX and the tile order over the XCDs (dispatch order, contiguous
blocks per XCD, or contiguous for the first RK stage only) at run time.  `tune` times
every (tile, SX, order) candidate on the GPU for one grid shape with HIP events
(libtbec's tbec_time_stages) and records the fastest in trapbec_amd/tuning.json;
`lookup` returns the recorded choice, else a default tile with an automatic segment
length (seg 0).  `tune_first_stage` then times a tiling of its own for the first RK
stage of real-time steps (the stage that reads one array) on top of that choice, and
`lookup_first_stage` returns it.
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
DEFAULT = {32: (2, 0), 64: (2, 0)}      # seg 0: libtbec picks it from the grid size
SEGS = (4, 8, 16, 32, 64, 128, 256)


def _key(shape, prec, sum_mode=None, w_stored=False, rot=False):
    if sum_mode is None:
        sum_mode = 0 if prec == 64 else 2
    return (f"{shape[0]}x{shape[1]}x{shape[2]}:f{prec}:s{sum_mode}" +
            (":ws" if w_stored else "") + (":rot" if rot else ""))


def _load():
    if os.path.exists(TABLE):
        with open(TABLE) as f:
            return json.load(f)
    return {}


def lookup(shape, prec, sum_mode=None, w_stored=False, rot=False, table=None):
    """(tile, seg, remap) recorded for this grid, precision, summation, w scheme and
    rotating frame (real time), in `table` (default trapbec_amd/tuning.json)."""
    if table is None:
        recs = _load()
    else:
        recs = json.load(open(table)) if os.path.exists(table) else {}
    t = recs.get(_key(shape, prec, sum_mode, w_stored, rot))
    if t:
        return int(t["tile"]), int(t["seg"]), int(t.get("remap", 1))
    return DEFAULT[prec] + (1,)


def tune(shape, prec, sum_mode=None, psi0=None, nstages=30, save=True, verbose=True,
         table=None, w_stored=False, problem=None, rot=False):
    """Time every (tile, seg) for this grid, precision, summation and w scheme (stages
    1-3 in turn); return the entry.  With `problem` (an initial.Problem) the stages of
    that problem are timed (rotation, imaginary time) and nothing is saved.  With
    `verbose` (the default) one line per candidate goes to stdout as it is timed."""
    import numpy as np
    from trapbec_amd import tbec
    if sum_mode is None:
        sum_mode = 0 if prec == 64 else 2
    if problem is not None:
        shape, save = tuple(problem.shape), False
        c = tbec.Condensate.from_problem(problem, prec=prec, sum_mode=sum_mode)
    else:
        c = tbec.Condensate(shape, (16.0, 16.0, 16.0), (256.0,) * 3, c0=0.0, gnl=1000.0,
                            dt=1e-4, prec=prec, sum_mode=sum_mode,
                            omega_z=-1.4 if rot else 0.0)
    c.set_w_storage(w_stored)
    if psi0 is None:
        psi0 = np.full(shape, 0.1 + 0.05j)
    c.upload(psi0)
    results = []
    for remap in (1, 0, 2):
        c.set_xcd_remap(remap)
        for tile in range(tbec.NUM_TILES):
            for seg in SEGS:
                c.set_tuning(tile, seg)
                c.time_stages(3)
                ms = c.time_stages(nstages)
                results.append((ms, tile, seg, remap))
                if verbose:
                    print(f"remap {remap} tile {tile} seg {seg}: {ms * 1e3:.1f} us/stage",
                          flush=True)
    c.close()
    ms, tile, seg, remap = min(results)
    entry = {"tile": tile, "seg": seg, "remap": remap, "us_per_stage": ms * 1e3,
             "all": [[t, s, r, round(m * 1e3, 2)] for m, t, s, r in sorted(results)]}
    if save:
        path = table or TABLE
        t = _load() if path == TABLE else (json.load(open(path)) if os.path.exists(path) else {})
        t[_key(shape, prec, sum_mode, w_stored, rot)] = entry
        with open(path, "w") as f:
            json.dump(t, f, indent=1, sort_keys=True)
    return entry


def lookup_first_stage(shape, prec, sum_mode=None, w_stored=False, rot=False, table=None):
    """(tile, seg) recorded for the first RK stage of real-time steps, or (-1, 0): the
    stage runs with the lookup() tiling."""
    if table is None:
        recs = _load()
    else:
        recs = json.load(open(table)) if os.path.exists(table) else {}
    t = recs.get(_key(shape, prec, sum_mode, w_stored, rot), {}).get("first_stage")
    if t:
        return int(t["tile"]), int(t["seg"])
    return -1, 0


def tune_first_stage(shape, prec, sum_mode=None, nstages=30, save=True, verbose=True,
                     table=None, w_stored=False, rot=False, reps=2):
    """With the recorded (tile, seg, remap) of this grid for the other stages, time every
    (tile, seg) for the first RK stage of real-time steps, and none (-1: the same tiling
    for all three), as the mean stage time of whole steps (tbec_time_stages over a
    multiple of 3 stages), `reps` rounds over the candidates, the fastest mean kept.
    Records it under "first_stage" of the grid's entry; returns that record."""
    import numpy as np
    from trapbec_amd import tbec
    if sum_mode is None:
        sum_mode = 0 if prec == 64 else 2
    if nstages % 3:
        nstages += 3 - nstages % 3
    tile, seg, remap = lookup(shape, prec, sum_mode, w_stored, rot, table)
    c = tbec.Condensate(shape, (16.0, 16.0, 16.0), (256.0,) * 3, c0=0.0, gnl=1000.0,
                        dt=1e-4, prec=prec, sum_mode=sum_mode,
                        omega_z=-1.4 if rot else 0.0)
    c.set_w_storage(w_stored)
    c.upload(np.full(shape, 0.1 + 0.05j))
    c.set_tuning(tile, seg)
    c.set_xcd_remap(remap)
    cands = [(-1, 0)] + [(t, sg) for t in range(tbec.NUM_TILES) for sg in SEGS]
    times = {k: [] for k in cands}
    for _ in range(3):                  # bring the clock and caches to their steady state
        c.time_stages(nstages)
    for _ in range(reps):
        for t1, s1 in cands:
            c.set_first_stage_tuning(t1, s1)
            c.time_stages(3)
            ms = c.time_stages(nstages)
            times[(t1, s1)].append(ms)
            if verbose:
                print(f"first stage tile {t1} seg {s1}: {ms * 1e3:.1f} us/stage", flush=True)
    c.close()
    res = sorted((float(np.mean(v)), k) for k, v in times.items())
    ms, (t1, s1) = res[0]
    rec = {"tile": t1, "seg": s1, "us_per_stage": ms * 1e3,
           "us_per_stage_same_tiling": float(np.mean(times[(-1, 0)])) * 1e3,
           "all": [[k[0], k[1], round(m * 1e3, 2)] for m, k in res]}
    if save:
        path = table or TABLE
        recs = _load() if path == TABLE else (json.load(open(path)) if os.path.exists(path)
                                              else {})
        key = _key(shape, prec, sum_mode, w_stored, rot)
        recs.setdefault(key, {"tile": tile, "seg": seg, "remap": remap})["first_stage"] = rec
        with open(path, "w") as f:
            json.dump(recs, f, indent=1, sort_keys=True)
    return rec


def tune_batch(problems, prec=32, sum_mode=None, steps=20, verbose=True):
    """Time every (tile, seg) for a batched ensemble: `steps` imaginary-time steps of all
    instances (one captured graph, constant states, no convergence stop).  Returns
    {"tile", "seg", "ms_per_step", "all"}; the caller applies it with Batch.set_tuning.
    With `verbose` (the default) one line per candidate goes to stdout as it is timed."""
    import time
    import numpy as np
    from trapbec_amd import tbec
    if steps & 1:
        steps += 1
    results = []
    with tbec.Batch(problems, prec=prec, sum_mode=sum_mode) as b:
        for i, pr in enumerate(problems):
            b.upload(i, np.full(pr.shape, 0.1 + 0.05j))
        for tile in range(tbec.NUM_TILES):
            for seg in SEGS:
                b.set_tuning(tile, seg)
                b.imag_solve(steps, 0.0, 1.0, steps)          # capture, warm up
                t0 = time.perf_counter()
                b.imag_solve(steps, 0.0, 1.0, steps)
                ms = (time.perf_counter() - t0) * 1e3 / steps
                results.append((ms, tile, seg))
                if verbose:
                    print(f"batch tile {tile} seg {seg}: {ms:.3f} ms/step", flush=True)
    ms, tile, seg = min(results)
    return {"tile": tile, "seg": seg, "ms_per_step": ms,
            "all": [[t, s, round(m, 4)] for m, t, s in sorted(results)]}


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="time every tile shape and x-segment length")
    ap.add_argument("--shape", type=int, nargs=3, default=[512, 512, 512])
    ap.add_argument("--prec", type=int, default=32)
    ap.add_argument("--sum", type=int, default=None)
    ap.add_argument("--w-stored", action="store_true", help="the stored-w RK3 scheme")
    ap.add_argument("--rot", action="store_true", help="rotating frame (Omega = 0.7)")
    ap.add_argument("--table", default=None, help="write here instead of tuning.json")
    ap.add_argument("--first-stage", action="store_true",
                    help="only the first RK stage's own tiling, on top of the recorded "
                         "choice for the grid")
    a = ap.parse_args()
    if a.first_stage:
        e = tune_first_stage(tuple(a.shape), a.prec, a.sum, verbose=True, table=a.table,
                             w_stored=a.w_stored, rot=a.rot)
    else:
        e = tune(tuple(a.shape), a.prec, a.sum, verbose=True, table=a.table,
                 w_stored=a.w_stored, rot=a.rot)
    print(json.dumps({"shape": a.shape, "prec": a.prec, "sum": a.sum,
                      "w_stored": a.w_stored, "rot": a.rot, **e}), flush=True)
