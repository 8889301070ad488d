"""The stopping verdict's two-level reduction, and a 4096^2 plane.

A checked launch leaves one partial residual sum per tile; on planes whose
largest problem has many tiles (CW_FOLD_TILES, 512 by default) a fold kernel
sums them once per problem and the next launch reads that sum, instead of
every workgroup summing all of its problem's partials. The fold sums in the
order a workgroup would, so the stopping decisions, the sweep counts and the
values are the same bits with the fold on or off; these tests force each way
on the same inputs and compare.

The 4096^2 plane (the three-field plane of the multi-field workload at 16 times
the area, about 2,700 to 3,200 tiles a field) is solved in FP64 and in FP32
smoothing with FP64 refinement to a relative residual of 1e-8; the oracle
recomputes each answer's residual, and the two precisions agree to 1e-6 of the
field's largest value.
"""
import math
import os
import time

import numpy as np
import pytest

from cellweave_amd import cases
from cellweave_amd.field import CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64, batch_relax
from helpers import bits_equal, gpu_from_case, oracle_from_case

pytestmark = pytest.mark.gpu

TOL = 1e-8


def say(msg):
    print(msg, flush=True)


def solve(fcs, precision, fold, monkeypatch, omega):
    """Solves fcs as one batch with the fold forced on (every plane) or off."""
    monkeypatch.setenv("CW_FOLD_TILES", "1" if fold else "1000000000")
    gs = [gpu_from_case(fc) for fc in fcs]
    t0 = time.perf_counter()
    st, _ = batch_relax(gs, criterion=CRIT_REL_RESIDUAL, tolerance=TOL, omega=omega,
                        precision=precision, max_sweeps=400_000)
    secs = time.perf_counter() - t0
    return [g.download() for g in gs], st, secs


def residual_by_oracle(fc, values):
    chk = oracle_from_case(fc)
    chk.values[:] = values
    return chk.rel_residual()[0]


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32_REFINE], ids=["fp64", "fp32refine"])
def test_fold_gives_the_same_bits_in_a_batch(precision, monkeypatch):
    """Six fields of four sizes in one launch sequence (several problems per
    launch, different tile counts): fold on and off, the same bits and stats."""
    say(f"case fold on/off, batch of six planes, precision {precision}")
    hist = cases.campaign_histories(count=3, lo=96, hi=400, seed=4600)
    fcs = [fc for h in hist for fc in h]
    a, sa, _ = solve(fcs, precision, True, monkeypatch, -1.0)
    b, sb, _ = solve(fcs, precision, False, monkeypatch, -1.0)
    for fc, x, y, p, q in zip(fcs, a, b, sa, sb):
        assert p["status"] == 0 and q["status"] == 0, (fc.name, p, q)
        assert p["sweeps"] == q["sweeps"] and p["outer_iters"] == q["outer_iters"], (p, q)
        assert p["rel_residual"] == q["rel_residual"], (p, q)
        assert bits_equal(x, y), fc.name
        assert residual_by_oracle(fc, x) <= TOL


def test_plane_4096(monkeypatch):
    say("case 4096^2 plane, oxygen: FP64 fold on and off, FP32+refine")
    fc = cases.multifield_plane(n=4096)[0]
    omega = 2.0 / (1.0 + math.sin(math.pi / 4096))
    (v64,), (s64,), t64 = solve([fc], PREC_FP64, True, monkeypatch, omega)
    say(f"  fp64 (fold): {t64:.3f} s, {s64['sweeps']} sweeps, rel {s64['rel_residual']:.3e}")
    (w64,), (u64,), u_t = solve([fc], PREC_FP64, False, monkeypatch, omega)
    say(f"  fp64 (no fold): {u_t:.3f} s, {u64['sweeps']} sweeps")
    assert s64["status"] == 0 and u64["status"] == 0
    assert s64["sweeps"] == u64["sweeps"] and bits_equal(v64, w64)
    monkeypatch.delenv("CW_FOLD_TILES")
    gs = [gpu_from_case(fc)]
    t0 = time.perf_counter()
    (s32,), _ = batch_relax(gs, criterion=CRIT_REL_RESIDUAL, tolerance=TOL, omega=omega,
                            precision=PREC_FP32_REFINE, max_sweeps=400_000)
    t32 = time.perf_counter() - t0
    v32 = gs[0].download()
    say(f"  fp32+refine: {t32:.3f} s, {s32['sweeps']} smoother sweeps, {s32['outer_iters']} "
        f"outer, rel {s32['rel_residual']:.3e}")
    assert s32["status"] == 0
    r64 = residual_by_oracle(fc, v64)
    r32 = residual_by_oracle(fc, v32)
    say(f"  oracle residuals: fp64 {r64:.3e}, fp32+refine {r32:.3e}")
    assert r64 <= TOL and r32 <= TOL
    scale = max(float(np.max(np.abs(v64))), 1e-30)
    diff = float(np.max(np.abs(v32 - v64))) / scale
    say(f"  fp32+refine against fp64: max |diff| / max|c| = {diff:.2e}")
    assert diff <= 1e-6 and (v32 >= 0).all()
