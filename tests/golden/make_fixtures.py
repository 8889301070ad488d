"""Writes the small golden fixtures and their SHA-256 sums.

Each fixture is one of the five workloads at a small size (planes 64^2 or
less, volumes 16^3), built by the seeded generators of cellweave_amd.cases,
with the oracle's FP64 answers:
  disc64        the 200^2 oxygen disc at 64^2 (omega 1.9), plus a lattice with
                obstacle ids and ids past the table, its per-cell gathers and
                the per-cell means of the answer;
  multifield64  the three-field 1024^2 plane at 64^2 (the reference's omega);
  spheroid16    the 256^3 spheroid at 16^3 (the reference's omega);
  campaign      two campaign histories (edges 40..64), two fields each;
  adjoint16     the 192^3 adjoint spheroid at 16^3: the solved field, the
                target, and the oracle's adjoint gradients.
For every field: the answer of the reference's stopping rule (largest change
below 1e-8, ScalarField::relax, pde.rs:245-288) and of the relative residual
1e-8 tested every sweep, each with its sweep count.

Run from the repository root:  python tests/golden/make_fixtures.py
The fixtures are data (inputs and expected outputs, numpy .npz without
pickles); tests/test_golden.py checks the sums, that the generators still make
the same inputs, that the oracle still gives the same outputs, and the
outputs against independent numpy restatements of the residual, the gathers
and the means; tests/test_gpu_golden.py runs the HIP path on them.
"""
from __future__ import annotations

import hashlib
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from cellweave_amd import cases  # noqa: E402
from oracle.orf import OracleField, optimal_omega  # noqa: E402

TOL = 1e-8
MAX_SWEEPS = 1_000_000
OBSTACLE = 0xFFFFFFFF


def fixture_cases():
    """name -> (field cases, omega rule): 'ref' (the reference's rule), a
    number, or 'optimal' (2 / (1 + sin(pi / longest edge)))."""
    return {
        "disc64": (cases.oxygen_disc_200(n=64), 1.9),
        "multifield64": (cases.multifield_plane(n=64), "ref"),
        "spheroid16": (cases.spheroid(n=16, ncells=40, radius=6.5, necrotic=2.5), "ref"),
        "campaign": ([fc for h in cases.campaign_histories(count=2, lo=40, hi=64) for fc in h],
                     "optimal"),
        "adjoint16": ([cases.adjoint_spheroid(n=16, ncells=30, radius=6.0, necrotic=2.0)[0]],
                      "optimal"),
    }


def oracle_of(fc):
    ref = OracleField(fc.w, fc.h, fc.d, fc.diff, fc.decay, fc.initial)
    ref.values[:] = fc.initial_values()
    ref.copy_next_from_values()
    ref.uptake[:] = fc.uptake
    ref.production[:] = fc.production
    ref.held[:] = fc.held
    if fc.source:
        ref.set_source(fc.source)
    return ref


def omega_of(fc, rule):
    if rule == "ref":
        return float(oracle_of(fc).reference_omega())
    if rule == "optimal":
        return optimal_omega(max(fc.w, fc.h, fc.d))
    return float(rule)


def coupling_lattice(fc, seed):
    """The field's lattice with a few obstacle voxels and ids past the table,
    per-cell uptake and production tables and a consumes flag per cell."""
    rng = np.random.default_rng(seed)
    grid = fc.grid.astype(np.uint32).copy()
    ncells = int(grid.max()) + 1
    cells = np.flatnonzero(grid != 0)
    pick = rng.choice(cells, size=24, replace=False)
    grid[pick[:12]] = OBSTACLE
    grid[pick[12:]] = ncells + rng.integers(0, 5, 12)
    up = np.concatenate([[0.0], rng.uniform(0.5e-3, 1.5e-3, ncells - 1)])
    prod = np.where(rng.random(ncells) < 0.3, rng.uniform(0.0, 1e-4, ncells), 0.0)
    prod[0] = 0.0
    consumes = (rng.random(ncells) < 0.8).astype(np.uint8)
    return grid, ncells, up, prod, consumes


def solve_field(fc, omega):
    """The oracle's answers under both stopping rules."""
    out = {}
    ref = oracle_of(fc)
    out["maxchange_sweeps"] = ref.relax(TOL, MAX_SWEEPS, omega=omega, threads=1)
    out["maxchange_values"] = ref.values.copy()
    ref = oracle_of(fc)
    out["residual_sweeps"] = ref.relax_residual(TOL, MAX_SWEEPS, omega=omega, threads=1)
    out["residual_values"] = ref.values.copy()
    out["residual_rel"] = ref.rel_residual()[0]
    return out


def build(name):
    fcs, rule = fixture_cases()[name]
    arrays = {"count": np.int64(len(fcs))}
    for k, fc in enumerate(fcs):
        p = f"f{k}_"
        omega = omega_of(fc, rule)
        arrays.update({
            p + "name": np.array(fc.name), p + "dims": np.array([fc.w, fc.h, fc.d], np.int64),
            p + "diff": np.float64(fc.diff), p + "decay": np.float64(fc.decay),
            p + "initial": np.float64(fc.initial), p + "hold_value": np.float64(fc.hold_value),
            p + "source": np.float64(fc.source), p + "omega": np.float64(omega),
            p + "held": fc.held.astype(np.uint8), p + "uptake": fc.uptake,
            p + "production": fc.production, p + "values0": fc.initial_values(),
        })
        for key, v in solve_field(fc, omega).items():
            arrays[p + key] = np.asarray(v)
    if name == "disc64":
        fc = fcs[0]
        grid, ncells, up, prod, consumes = coupling_lattice(fc, 1001)
        ref = oracle_of(fc)
        ref.apply_uptake_per_cell(grid, consumes, up)
        ref.apply_production(grid, prod)
        gathered_up, gathered_prod = ref.uptake.copy(), ref.production.copy()
        ref.values[:] = arrays["f0_maxchange_values"]
        arrays.update({
            "grid": grid, "ncells": np.int64(ncells), "cell_uptake": up,
            "cell_production": prod, "consumes": consumes, "gathered_uptake": gathered_up,
            "gathered_production": gathered_prod,
            "means": ref.mean_per_cell(grid, ncells),
        })
    if name == "adjoint16":
        fc = fcs[0]
        _, params = cases.adjoint_spheroid(n=16, ncells=30, radius=6.0, necrotic=2.0)
        omega = float(arrays["f0_omega"])
        # the target: the FP64 solve with D and each cell's uptake perturbed
        tgt = cases.FieldCase(**{**fc.__dict__, "diff": fc.diff * params["diff_scale"],
                                 "uptake": cases.gather(fc.grid, fc.per_cell_uptake *
                                                        params["uptake_scale"])})
        t = oracle_of(tgt)
        t.relax_residual(1e-14, MAX_SWEEPS, omega=omega)
        base = oracle_of(fc)
        base.relax_residual(1e-14, MAX_SWEEPS, omega=omega)
        ncells = int(fc.grid.max()) + 1
        adj = base.adjoint(t.values, omega=omega, rel_tol=1e-13, grid=fc.grid, ncells=ncells)
        arrays.update({
            "grid": fc.grid.astype(np.uint32), "ncells": np.int64(ncells),
            "solved": base.values.copy(), "target": t.values.copy(),
            "dl_dd": np.float64(adj["dD"]), "dl_dg": np.float64(adj["dg"]), "dl_du": adj["du"],
        })
    return arrays


NAMES = ("disc64", "multifield64", "spheroid16", "campaign", "adjoint16")


def sha256(path):
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def main():
    lines = []
    for name in NAMES:
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **build(name))
        lines.append(f"{sha256(path)}  {name}.npz")
        print(lines[-1], flush=True)
    with open(os.path.join(HERE, "SHA256SUMS"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
