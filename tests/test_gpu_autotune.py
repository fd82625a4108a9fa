"""The autotuner (option "autotune"): the launch shapes it picks change no
result beyond rounding, a plan built with the picked shapes set explicitly is
identical bit for bit, and a tuning file written by one context is read by the
next, which then times nothing."""
import math

import numpy as np
import pytest

import BoostFractor as B
from BoostFractor import workloads as W
from helpers import DISKR, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("mixed", 1e-4)])
def test_dancer_autotune_shapes(tmp_path, prec, tol):
    coords = W.grid_coords(64)
    sb = W.booster(coords, ndisk=3, seed=1701)
    freqs = list(np.linspace(19.6e9, 20.4e9, 7))
    tune_file = tmp_path / "tune.txt"
    ctx = B.Context()
    try:
        ctx.set_tuning_file(str(tune_file))
        ctx.set_option("autotune", 1)
        print(f"dancer autotune {prec}: 64x64, 3 disks, 7 freqs", flush=True)
        plan = B.DancerPlan(0, 40, sb, coords, freqs, diskR=DISKR, precision=prec, ctx=ctx)
        plan.run()
        E, p = plan.fetch()
        plan.close()
        tuned, ms = ctx.tuning_report()
        keys = [k for k in tuned if k.startswith("dancer")]
        assert len(keys) == 1 and ms > 0, (tuned, ms)
        fb, tx, rows = tuned[keys[0]]
        assert 1 <= fb <= len(freqs) and tx >= 0 and rows >= 0
        # the same shapes, set explicitly: bit for bit
        nb = math.ceil(len(freqs) / fb)
        with ctx.options(autotune=0, col_tx=tx, rows=rows):
            ref = B.DancerPlan(0, 40, sb, coords, freqs, diskR=DISKR, precision=prec, ctx=ctx,
                               freq_batch=math.ceil(len(freqs) / nb))
            ref.run()
            E2, p2 = ref.fetch()
            ref.close()
        assert np.array_equal(E, E2) and np.array_equal(p, p2)
        # the heuristic shapes: within rounding
        with ctx.options(autotune=0):
            heur = B.DancerPlan(0, 40, sb, coords, freqs, diskR=DISKR, precision=prec, ctx=ctx)
            heur.run()
            E3, _ = heur.fetch()
            heur.close()
        assert rel(E, E3) < tol, rel(E, E3)
        # a second context reads the file and times nothing
        lines = tune_file.read_text().splitlines()
        assert len(lines) == 1 and lines[0].split("\t")[0] == keys[0]
        ctx2 = B.Context()
        try:
            ctx2.set_tuning_file(str(tune_file))
            assert ctx2.tuning_report() == ({keys[0]: [fb, tx, rows]}, 0.0)
            ctx2.set_option("autotune", 1)
            plan = B.DancerPlan(0, 40, sb, coords, freqs, diskR=DISKR, precision=prec, ctx=ctx2)
            plan.run()
            E4, _ = plan.fetch()
            plan.close()
            assert ctx2.tuning_report()[1] == 0.0
            assert np.array_equal(E4, E)
        finally:
            ctx2.close()
    finally:
        ctx.close()


def test_sweep_autotune_parseval_blocks():
    """The sweep plan with the autotuner on (Parseval product column blocks timed
    on the first large flush) equals the plan with either width set explicitly;
    both widths sum every matrix entry in the same order, so all three agree
    bit for bit.  The spacing series is off here, so that every air gap goes
    through the direct product that is tuned."""
    coords = W.grid_coords(64)
    geoms, set_of = W.sweep_geometries(coords, 12)
    sets = [B.SeedModes(coords, ThreeDim=True, Mmax=M, Lmax=L_, diskR=DISKR) for M, L_ in W.SWEEP_MODESETS]
    emits = [B.axion_induced_modes(coords, md, diskR=DISKR) for md in sets]
    freqs = np.linspace(19e9, 21e9, 100)
    nprob = sum(len(g.distance) // 2 for g in geoms) * len(freqs)
    assert nprob >= 2 * 64 * 256  # enough flat air regions for a timed flush
    out = {}
    ctx = B.Context()
    try:
        for name, kw in (("auto", dict(autotune=1, parseval_series=0)), ("2", dict(pg_wn=2, parseval_series=0)),
                         ("4", dict(pg_wn=4, parseval_series=0))):
            print(f"sweep autotune: 12 geometries, 64x64, 100 freqs, options {kw}", flush=True)
            with ctx.options(**kw):
                plan = B.TransformerBatchPlan(coords, sets, freqs, diskR=DISKR, emits=emits, ctx=ctx)
                plan.run([g.copy() for g in geoms], mode_set_of=set_of)
                out[name] = plan.fetch()
                plan.close()
        tuned, ms = ctx.tuning_report()
        keys = [k for k in tuned if k.startswith("pgemm")]
        assert keys and ms > 0 and all(tuned[k][0] in (2, 4) for k in keys), tuned
    finally:
        ctx.close()
    for c in range(len(geoms)):
        assert np.array_equal(out["auto"][c], out["2"][c]), c
        assert np.array_equal(out["auto"][c], out["4"][c]), c
