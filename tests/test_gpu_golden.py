"""The HIP path through the C ABI on the committed golden fixtures
(tests/golden/*.npz): every field of the five small workloads solved under
the reference's stopping rule (largest change below 1e-8) and under the
relative residual tested every sweep gives the fixture's sweep count and
values bit for bit; the device gathers and per-cell means of the disc
fixture, and the adjoint gradients of the 16^3 spheroid, against the
fixture's outputs; FP32 smoothing with FP64 refinement reaches 1e-8 on every
field and lands within 1e-6 of the fixture's FP64 answer."""
import os

import numpy as np
import pytest

from cellweave_amd.field import (CRIT_MAX_CHANGE, CRIT_REL_RESIDUAL, PREC_FP32_REFINE, Field,
                                 Lattice)
from helpers import bits_equal

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("disc64", "multifield64", "spheroid16", "campaign", "adjoint16")
TOL = 1e-8


def say(msg):
    print(msg, flush=True)


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def fields(z):
    for k in range(int(z["count"])):
        p = f"f{k}_"
        yield k, {key[len(p):]: z[key] for key in z.files if key.startswith(p)}


def device_field(f, values=None):
    w, h, d = (int(v) for v in f["dims"])
    g = Field(w, h, d, float(f["diff"]), float(f["decay"]), float(f["initial"]))
    if float(f["source"]):
        g.set_source(float(f["source"]))
    g.upload(values=f["values0"] if values is None else values, held=f["held"],
             uptake=f["uptake"], production=f["production"])
    return g


@pytest.mark.parametrize("name", NAMES)
def test_max_change_rule_bitwise(name):
    z = load(name)
    for k, f in fields(z):
        say(f"case golden {name} field {k} max-change rule")
        g = device_field(f)
        st = g.relax_ex(omega=float(f["omega"]), criterion=CRIT_MAX_CHANGE, tolerance=TOL,
                        max_sweeps=1_000_000)
        assert st["sweeps"] == int(f["maxchange_sweeps"]), (name, k, st)
        assert bits_equal(g.download(), f["maxchange_values"]), (name, k)


@pytest.mark.parametrize("name", ["multifield64", "spheroid16"])
def test_reference_entry_point_bitwise(name):
    """cw_field_relax (ScalarField::relax: the reference's omega and rule)."""
    z = load(name)
    for k, f in fields(z):
        say(f"case golden {name} field {k} cw_field_relax")
        g = device_field(f)
        assert g.relax(TOL, 1_000_000) == int(f["maxchange_sweeps"])
        assert bits_equal(g.download(), f["maxchange_values"]), (name, k)


@pytest.mark.parametrize("name", ["disc64", "multifield64", "campaign"])
def test_residual_rule_bitwise(name):
    """Planes, one sweep a launch, the residual of every launch's input tested."""
    z = load(name)
    for k, f in fields(z):
        say(f"case golden {name} field {k} relative residual")
        g = device_field(f)
        st = g.relax_ex(omega=float(f["omega"]), criterion=CRIT_REL_RESIDUAL, tolerance=TOL,
                        max_sweeps=1_000_000, sweeps_per_pass=1, check_every=1)
        assert st["sweeps"] == int(f["residual_sweeps"]), (name, k, st)
        assert bits_equal(g.download(), f["residual_values"]), (name, k)
        assert abs(st["rel_residual"] - float(f["residual_rel"])) <= 1e-12 * float(f["residual_rel"])


@pytest.mark.parametrize("name", NAMES)
def test_fp32_refine_against_the_fixture(name):
    z = load(name)
    for k, f in fields(z):
        say(f"case golden {name} field {k} fp32+refine")
        g = device_field(f)
        st = g.relax_ex(omega=float(f["omega"]), criterion=CRIT_REL_RESIDUAL, tolerance=TOL,
                        max_sweeps=1_000_000, precision=PREC_FP32_REFINE)
        assert st["status"] == 0 and st["rel_residual"] <= TOL, (name, k, st)
        v, want = g.download(), f["residual_values"]
        assert np.max(np.abs(v - want)) <= 1e-6 * max(1e-30, float(np.max(np.abs(want))))
        assert (v >= 0).all()


def test_disc_gathers_and_means():
    say("case golden disc64 gathers and means")
    z = load("disc64")
    (_, f), = fields(z)
    lat = Lattice(64, 64, 1)
    lat.upload(z["grid"])
    g = device_field(f)
    g.apply_uptake_per_cell(lat, z["cell_uptake"], consumes=z["consumes"])
    g.apply_production(lat, z["cell_production"])
    _, up, pr = g.state()
    assert bits_equal(up, z["gathered_uptake"]) and bits_equal(pr, z["gathered_production"])
    g.upload(values=f["maxchange_values"])
    assert bits_equal(g.mean_per_cell(lat, int(z["ncells"])), z["means"])


def test_adjoint_gradients():
    say("case golden adjoint16")
    z = load("adjoint16")
    (_, f), = fields(z)
    g = device_field(f, values=z["solved"])
    got = g.adjoint(z["target"], grid=z["grid"], ncells=int(z["ncells"]), tolerance=1e-13,
                    omega=float(f["omega"]))
    assert abs(got["dl_dd"] - float(z["dl_dd"])) <= 1e-9 * abs(float(z["dl_dd"]))
    assert abs(got["dl_dg"] - float(z["dl_dg"])) <= 1e-9 * abs(float(z["dl_dg"]))
    du = z["dl_du"]
    assert np.max(np.abs(got["du"] - du)) <= 1e-9 * np.max(np.abs(du))
