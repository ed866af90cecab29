"""The autotuner on the GPU: every (tile, segment, XCD order) candidate is timed and the
recorded choice is the fastest of its own sweep; the choice changes speed only (the
tuned and the default launch give the same psi); the batch tuner likewise."""
import numpy as np
import pytest

from trapbec_amd import autotune, initial, tbec

pytestmark = pytest.mark.gpu


def test_tune_records_the_fastest_candidate(tmp_path):
    shape, prec = (40, 36, 48), 32
    table = str(tmp_path / "tuning.json")
    print(f"\n[case] autotune {shape} fp{prec}", flush=True)
    e = autotune.tune(shape, prec, nstages=6, table=table)
    assert len(e["all"]) == 3 * tbec.NUM_TILES * len(autotune.SEGS)
    best = min(a[3] for a in e["all"])
    assert round(e["us_per_stage"], 2) == best
    assert [e["tile"], e["seg"], e["remap"]] == e["all"][0][:3]
    assert autotune.lookup(shape, prec, table=table) == (e["tile"], e["seg"], e["remap"])
    # the recorded choice changes the launch shape, not the result
    pr = initial.Problem(shape, (8.0, 7.0, 9.0), (1.0, 1.1, 0.9), 50.0, 0.4, 1e-3)
    psi0 = initial.tf_profile(shape, pr.box, pr.omega, pr.g, norm=1.0)[0] * \
        initial.vortex_line_factor(shape, pr.box, 0.3, 0.1, 1, 0.4)
    out = []
    for tile, seg, remap in ((e["tile"], e["seg"], e["remap"]), (2, 0, 1)):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.set_tuning(tile, seg)
            c.set_xcd_remap(remap)
            c.upload(psi0)
            c.step(40)
            out.append(c.download())
    err = float(np.abs(out[0] - out[1]).max() / np.abs(out[1]).max())
    print(f"tuned {e['tile'], e['seg'], e['remap']} vs default: rel max {err:.2e}", flush=True)
    assert err < 2e-6


def test_tune_batch_picks_the_fastest_candidate():
    probs = [p for p, _ in initial.config_c4(seed=3, n_inst=3, dx=0.4, dtau=4e-3,
                                              grids=((24, 24, 24), (32, 20, 28)))]
    print(f"\n[case] tune_batch {[p.shape for p in probs]}", flush=True)
    e = autotune.tune_batch(probs, prec=32, steps=8)
    assert len(e["all"]) == tbec.NUM_TILES * len(autotune.SEGS)
    assert [e["tile"], e["seg"]] == e["all"][0][:2]
    assert round(e["ms_per_step"], 4) == min(a[2] for a in e["all"])


@pytest.mark.parametrize("prec", [32, 64])
def test_first_stage_tiling_changes_speed_not_results(prec):
    """The first RK stage run with each tile shape of its own (the other stages with the
    handle's tiling) gives the psi of the single tiling, on a ragged rotating grid."""
    shape = (40, 36, 52)
    pr = initial.Problem(shape, (8.0, 7.0, 9.0), (1.0, 1.1, 0.9), 50.0, 0.4, 1e-3)
    psi0 = initial.tf_profile(shape, pr.box, pr.omega, pr.g, norm=1.0)[0] * \
        initial.vortex_line_factor(shape, pr.box, 0.3, 0.1, 1, 0.4)
    print(f"\n[case] first-stage tiling {shape} fp{prec}", flush=True)

    def run(tile1, seg1):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.set_tuning(2, 0)
            c.set_first_stage_tuning(tile1, seg1)
            c.upload(psi0)
            c.step(3)           # eager
            c.step(32)          # graph replays
            return c.download()

    ref = run(-1, 0)
    tol = 2e-6 if prec == 32 else 1e-12
    for tile1 in range(tbec.NUM_TILES):
        for seg1 in (0, 16):
            got = run(tile1, seg1)
            err = float(np.abs(got - ref).max() / np.abs(ref).max())
            print(f"first stage tile {tile1} seg {seg1}: rel max {err:.2e}", flush=True)
            assert err < tol, (tile1, seg1, err)


@pytest.mark.parametrize("setup", ["table_solid", "periodic_advective"])
def test_first_stage_tiling_with_other_physics(setup):
    """The first stage's own tiling with the trap read from a table, boundary faces that a
    fill kernel refreshes before each stage, and the advective drift: the psi of the
    single tiling (FP32 centred sums, 40 x 36 x 52)."""
    shape = (40, 36, 52)
    pr = initial.Problem(shape, (8.0, 7.0, 9.0), (1.0, 1.1, 0.9), 50.0, 0.0, 1e-3)
    psi0 = initial.tf_profile(shape, pr.box, pr.omega, pr.g, norm=1.0)[0] * \
        initial.vortex_line_factor(shape, pr.box, 0.3, 0.1, 1, 0.4)
    print(f"\n[case] first-stage tiling {setup} {shape}", flush=True)

    def run(tile1):
        with tbec.Condensate.from_problem(pr, prec=32) as c:
            if setup == "table_solid":
                tab = c.potential(0.0)
                c.set_physics("table")
                c.set_potential_table(tab)
                c.set_boundaries("solid")
            else:
                c.set_physics("harmonic", advective_vx=0.3)
                c.set_boundaries("periodic", periodicz=True)
            c.set_tuning(2, 0)
            c.set_first_stage_tuning(tile1, 0)
            c.upload(psi0)
            c.step(3)
            c.step(32)
            return c.download()

    ref = run(-1)
    for tile1 in range(tbec.NUM_TILES):
        got = run(tile1)
        err = float(np.abs(got - ref).max() / np.abs(ref).max())
        print(f"first stage tile {tile1}: rel max {err:.2e}", flush=True)
        assert err < 2e-6, (tile1, err)


def test_first_stage_tuning_rejects_bad_arguments():
    pr = initial.Problem((16, 16, 16), (8.0, 8.0, 8.0), (1.0, 1.0, 1.0), 50.0, 0.0, 1e-3)
    with tbec.Condensate.from_problem(pr, prec=32) as c:
        for tile, seg in ((tbec.NUM_TILES, 0), (-2, 0), (0, -1), (0, 5000)):
            with pytest.raises(tbec.TbecError):
                c.set_first_stage_tuning(tile, seg)
        c.set_first_stage_tuning(-1, 0)


def test_tune_first_stage_records_the_fastest(tmp_path):
    shape, prec = (40, 36, 48), 32
    table = str(tmp_path / "tuning.json")
    print(f"\n[case] first-stage autotune {shape} fp{prec}", flush=True)
    rec = autotune.tune_first_stage(shape, prec, nstages=6, table=table, reps=1)
    assert len(rec["all"]) == 1 + tbec.NUM_TILES * len(autotune.SEGS)
    assert [rec["tile"], rec["seg"]] == rec["all"][0][:2]
    assert round(rec["us_per_stage"], 2) == min(a[2] for a in rec["all"])
    assert autotune.lookup_first_stage(shape, prec, table=table) == (rec["tile"], rec["seg"])
