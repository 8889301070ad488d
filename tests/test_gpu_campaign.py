"""The device-driven batch of planes (cw_campaign_relax): per-field phases on
the device, one captured graph for every size class.

FP64 must give cw_batch_relax's sweeps and bits at the same tile shape and
test cadence (the same sweep kernel, the same fixed-order residual sums). FP32
refinement must reach the tolerance by the oracle's own FP64 residual and
agree with FP64; budgets are hard limits."""
import ctypes as C

import numpy as np
import pytest

from cellweave_amd import _abi, campaign, cases
from cellweave_amd.field import (CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64, Field,
                                 batch_relax, campaign_relax)
from helpers import bits_equal, gpu_from_case, oracle_from_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def small_campaign(count=10, lo=48, hi=300, seed=4100):
    hist = cases.campaign_histories(seed=seed, count=count, lo=lo, hi=hi)
    return [fc for h in hist for fc in h]


def groups_of(flat):
    pl = campaign.plan(flat, PREC_FP64, table={})
    groups = [0] * len(flat)
    for g, (_cls, idx, _shape) in enumerate(pl):
        for k in idx:
            groups[k] = g
    return groups, len(pl)


@pytest.mark.parametrize("shape", [8, 9])
def test_fp64_device_driven_equals_host_driven(shape):
    say(f"case device-driven FP64, shape {shape}")
    flat = small_campaign()
    groups, ng = groups_of(flat)
    dev = [gpu_from_case(fc) for fc in flat]
    st_d, _ = campaign_relax(dev, groups, [shape] * ng, criterion=CRIT_REL_RESIDUAL,
                             tolerance=1e-8, omega=-1.0, precision=PREC_FP64)
    for g in range(ng):
        idx = [k for k in range(len(flat)) if groups[k] == g]
        host = [gpu_from_case(flat[k]) for k in idx]
        st_h, _ = batch_relax(host, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=-1.0,
                              precision=PREC_FP64, tile_shape=shape, use_graph=1)
        for k, sh, hf in zip(idx, st_h, host):
            sd = st_d[k]
            assert sd["status"] == 0 and sd["tile_shape"] == shape
            assert sd["sweeps"] == sh["sweeps"], (k, sd["sweeps"], sh["sweeps"])
            assert sd["omega"] == sh["omega"]
            assert bits_equal(dev[k].download(), hf.download()), k
            # ||b_eff||^2 is summed per wave here and, for batches of up to
            # four fields, by whole workgroups there: equal to rounding
            assert abs(sd["rel_residual"] - sh["rel_residual"]) <= 1e-12 * sh["rel_residual"]


@pytest.mark.parametrize("shape", [15, 19, 23])
def test_fp32_device_driven_converges(shape):
    say(f"case device-driven FP32, shape {shape}")
    flat = small_campaign(seed=4200)
    groups, ng = groups_of(flat)
    g32 = [gpu_from_case(fc) for fc in flat]
    g64 = [gpu_from_case(fc) for fc in flat]
    st32, _ = campaign_relax(g32, groups, [shape] * ng, criterion=CRIT_REL_RESIDUAL,
                             tolerance=1e-8, omega=-1.0, precision=PREC_FP32_REFINE)
    st64, _ = campaign_relax(g64, groups, None, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8,
                             omega=-1.0, precision=PREC_FP64)
    for k, fc in enumerate(flat):
        assert st32[k]["status"] == 0 and st32[k]["rel_residual"] <= 1e-8
        assert st32[k]["outer_iters"] >= 1
        v32, v64 = g32[k].download(), g64[k].download()
        chk = oracle_from_case(fc)
        chk.values[:] = v32
        assert chk.rel_residual()[0] <= 1e-8, k
        assert (v32 >= 0).all()
        assert np.max(np.abs(v32 - v64)) <= 1e-6 * max(float(np.max(np.abs(v64))), 1e-30)


def test_repeat_gives_the_same_bits():
    say("case device-driven repeat")
    flat = small_campaign(count=6, seed=4300)
    groups, ng = groups_of(flat)
    for prec in (PREC_FP64, PREC_FP32_REFINE):
        fs = [gpu_from_case(fc) for fc in flat]
        st_a, _ = campaign_relax(fs, groups, None, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8,
                                 omega=-1.0, precision=prec)
        va = [f.download() for f in fs]
        for f, fc in zip(fs, flat):
            f.upload(values=fc.initial_values())
        st_b, _ = campaign_relax(fs, groups, None, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8,
                                 omega=-1.0, precision=prec)
        for a, b, x, f in zip(st_a, st_b, va, fs):
            assert a["sweeps"] == b["sweeps"]
            assert bits_equal(x, f.download())


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32_REFINE])
def test_budget_is_a_hard_limit(prec):
    """Fields with different sizes under one max_sweeps: none runs past it,
    every one that did not converge is reported, the call fails with the
    reference's message."""
    say(f"case device-driven budget prec={prec}")
    flat = small_campaign(count=6, lo=100, hi=300, seed=4400)
    groups, ng = groups_of(flat)
    fs = [gpu_from_case(fc) for fc in flat]
    st, rc = campaign_relax(fs, groups, None, check=False, criterion=CRIT_REL_RESIDUAL,
                            tolerance=1e-8, omega=-1.0, precision=prec, max_sweeps=100)
    assert rc == _abi.CW_ERR_NOT_CONVERGED
    assert "did not settle within 100 sweeps" in _abi.last_error()
    for s in st:
        assert s["sweeps"] <= 100
        assert s["status"] == _abi.CW_ERR_NOT_CONVERGED


def test_nothing_free_and_refusals():
    say("case device-driven edge cases")
    (fc,) = cases.oxygen_disc_200()
    a = gpu_from_case(fc)
    held = Field(40, 30, 1, 0.2, 0.0, 1.0)
    held.upload(held=np.ones(40 * 30, np.uint8))
    st, _ = campaign_relax([a, held], [0, 0], None, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8,
                           omega=1.9, precision=PREC_FP64)
    assert st[1]["sweeps"] == 0 and st[1]["status"] == 0
    assert st[0]["status"] == 0
    chk = oracle_from_case(fc)
    chk.values[:] = a.download()
    assert chk.rel_residual()[0] <= 1e-8
    # shapes the batch is not compiled for, volumes, the max-change rule
    _, rc = campaign_relax([a], [0], [12], check=False, criterion=CRIT_REL_RESIDUAL,
                           precision=PREC_FP32_REFINE)
    assert rc == _abi.CW_ERR_ARG and "tile shape 12" in _abi.last_error()
    vol = Field(8, 8, 8, 0.2, 0.0, 1.0)
    _, rc = campaign_relax([vol], [0], None, check=False, criterion=CRIT_REL_RESIDUAL)
    assert rc == _abi.CW_ERR_ARG and "planes only" in _abi.last_error()
    _, rc = campaign_relax([a], [0], None, check=False)
    assert rc == _abi.CW_ERR_ARG and "relative residual" in _abi.last_error()
