"""Shared builders for the tests: the same inputs for the oracle and the GPU."""
from __future__ import annotations

import numpy as np

from oracle.orf import OracleField

D = 0.2  # pde.rs:758, the disc tests' diffusion coefficient


def outside_disc(size, radius):
    """pde.rs:748-755: everything outside a disc of `radius` centred in a box of `size`."""
    c = float(size // 2)
    yy, xx = np.mgrid[0:size, 0:size]
    return ((xx - c) ** 2 + (yy - c) ** 2 > float(radius) ** 2).reshape(-1)


def bathed_disc_inputs(size, radius, uptake):
    """pde.rs:762-775 up to the relax: uptake inside the disc, bath at 1 outside."""
    out = outside_disc(size, radius)
    ref = OracleField(size, size, 1, D, 0.0, 1.0)
    up = np.where(out, 0.0, uptake)
    ref.uptake[:] = up
    ref.hold_bath_at(1.0, out.astype(np.uint8))
    return ref


def oracle_from_case(fc):
    ref = OracleField(fc.w, fc.h, fc.d, fc.diff, fc.decay, fc.initial)
    ref.values[:] = fc.initial_values()
    ref.copy_next_from_values()
    ref.uptake[:] = fc.uptake
    ref.production[:] = fc.production
    ref.held[:] = fc.held
    if fc.source:
        ref.set_source(fc.source)
    return ref


def gpu_from_oracle(ref):
    from cellweave_amd.field import Field

    g = Field(ref.w, ref.h, ref.d, ref.diff, ref.decay, 0.0)
    g.upload(values=ref.values.copy(), held=ref.held.copy(), uptake=ref.uptake.copy(),
             production=ref.production.copy())
    return g


def gpu_from_case(fc):
    from cellweave_amd.field import Field

    g = Field(fc.w, fc.h, fc.d, fc.diff, fc.decay, fc.initial)
    if fc.source:
        g.set_source(fc.source)
    g.upload(values=fc.initial_values(), held=fc.held, uptake=fc.uptake,
             production=fc.production)
    return g


def bits_equal(a, b):
    """Bitwise equality, with +0.0 and -0.0 taken as equal (see the floor)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64)) or (
        np.array_equal(a, b) and np.array_equal(np.isnan(a), np.isnan(b)))
