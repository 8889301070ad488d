"""The autotuner: cw_autotune ranks every sweep shape of a batch by its
launch time per sweep, and the campaign's table takes the fastest whole
solve among the best-ranked shapes of each size class."""
import numpy as np
import pytest

from cellweave_amd import campaign, cases
from cellweave_amd.field import (CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64, Field, autotune,
                                 batch_relax)
from helpers import gpu_from_case

pytestmark = pytest.mark.gpu

# shape tables of the library: plain FP64 planes, FP32 smoother planes, volumes
N_PLAIN, N_SMOOTH, N_VOLUME = 10, 28, 19
TWO_SWEEP_VOLUME = {12, 13, 14, 18}  # FP32 smoothing only


def say(msg):
    print(msg, flush=True)


def _ranked_ok(ranked, want):
    shapes = [s for s, _ in ranked]
    us = [u for _, u in ranked]
    assert sorted(shapes) == sorted(want)
    assert us == sorted(us)
    assert all(np.isfinite(u) and u > 0 for u in us)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32_REFINE])
def test_autotune_ranks_every_plane_shape(precision):
    say(f"case autotune 2 x 192^2 planes precision={precision}")
    fcs = cases.multifield_plane(n=192)[:2]
    gs = [gpu_from_case(fc) for fc in fcs]
    ranked = autotune(gs, launches=3, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=-1.0,
                      precision=precision)
    _ranked_ok(ranked, range(N_PLAIN if precision == PREC_FP64 else N_SMOOTH))
    # the fields still solve after timing, with the winner, to the tolerance
    for g, fc in zip(gs, fcs):
        g.fill(fc.initial)
    st, _ = batch_relax(gs, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=-1.0,
                        precision=precision, tile_shape=ranked[0][0])
    assert all(s["status"] == 0 and s["rel_residual"] <= 1e-8 for s in st)
    assert [s["tile_shape"] for s in st] == [ranked[0][0]] * 2


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32_REFINE])
def test_autotune_ranks_every_volume_shape(precision):
    say(f"case autotune 40^3 volume precision={precision}")
    g = Field(40, 40, 40, 0.5, 0.01, 1.0)
    g.set_source(1e-3)
    ranked = autotune([g], launches=2, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=-1.0,
                      precision=precision)
    want = [s for s in range(N_VOLUME)
            if precision == PREC_FP32_REFINE or s not in TWO_SWEEP_VOLUME]
    _ranked_ok(ranked, want)


def test_campaign_table_takes_fastest_confirmed_solve():
    say("case campaign tune, 6 histories")
    hist = cases.campaign_histories(count=6, lo=96, hi=320)
    flat = [fc for h in hist for fc in h]
    table = campaign.tune(flat, launches=2, confirm=3)
    classes = sorted(campaign.group(flat))
    for key, n in (("fp64", N_PLAIN), ("fp32", N_SMOOTH)):
        assert sorted(table[key]) == classes
        for cls in classes:
            timed = table["confirmed_ms"][f"{key} {cls}"]
            ranked = table["ranking"][f"{key} {cls}"]
            # the ranking covers the shapes the device-driven campaign runs;
            # the confirmed ones are its (up to) three best, timed in ms
            assert all(s in campaign.DEVICE_SHAPES[key] for s, _ in ranked)
            assert len(timed) == min(3, len(campaign.DEVICE_SHAPES[key]))
            assert sorted(s for s, _ in timed) == sorted(s for s, _ in ranked[:len(timed)])
            assert [t for _, t in timed] == sorted(t for _, t in timed)
            assert table[key][cls] == timed[0][0] and 0 <= timed[0][0] < n
    # the table drives a campaign solve to the tolerance
    fields = campaign.make_fields(flat)
    for key, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32_REFINE)):
        for g, fc in zip(fields, flat):
            g.fill(fc.initial)
        stats, _ = campaign.solve(fields, campaign.plan(flat, prec, table), precision=prec)
        assert all(s["status"] == 0 and s["rel_residual"] <= 1e-8 for s in stats)
