"""cw_field_relax_noalloc against cw_field_relax, through the C ABI.

The allocation-free relax runs the same kernels as cw_field_relax with its
launch plan in the library's static device scratch, so it must give the same
sweep count and the same bits. Two calls in a row, the second on a smaller
field, then a volume: every call reuses the one scratch array, which keeps the
previous call's contents.
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import bathed_disc_inputs, bits_equal, gpu_from_oracle

pytestmark = pytest.mark.gpu


def _abi():
    from cellweave_amd import _abi

    return _abi


def relax_both(ref, tol, max_sweeps):
    """(sweeps, values) from cw_field_relax and from cw_field_relax_noalloc,
    each on its own device copy of the same inputs."""
    ab = _abi()
    lib = ab.lib()
    out = []
    for name in ("cw_field_relax", "cw_field_relax_noalloc"):
        g = gpu_from_oracle(ref)
        n = C.c_uint64()
        rc = getattr(lib, name)(g._h, tol, max_sweeps, C.byref(n))
        assert rc == ab.CW_OK, (name, rc, ab.last_error())
        out.append((int(n.value), g.download()))
    return out


def ball_inputs(n, radius, uptake):
    """A consuming ball in a box whose faces are held at 1 (pde.rs:1203-1229)."""
    from oracle.orf import OracleField

    c = (n - 1) / 2.0
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:n]
    inside = ((xx - c) ** 2 + (yy - c) ** 2 + (zz - c) ** 2 <= radius * radius).reshape(-1)
    face = ((xx == 0) | (xx == n - 1) | (yy == 0) | (yy == n - 1) | (zz == 0) |
            (zz == n - 1)).reshape(-1)
    ref = OracleField(n, n, n, 0.2, 0.0, 1.0)
    ref.uptake[:] = np.where(inside, uptake, 0.0)
    ref.held[:] = face.astype(np.uint8)
    return ref


def test_noalloc_relax_matches_relax_on_consecutive_calls():
    cases = [bathed_disc_inputs(96, 30, 0.004), bathed_disc_inputs(30, 6, 0.004),
             ball_inputs(20, 6, 0.01)]
    for ref in cases:
        print(f"noalloc relax: {ref.w}x{ref.h}x{ref.d}", flush=True)
        (s0, v0), (s1, v1) = relax_both(ref, 1e-10, 200_000)
        assert s0 == s1, (ref.w, ref.h, ref.d, s0, s1)
        assert bits_equal(v0, v1), (ref.w, ref.h, ref.d, float(np.max(np.abs(v0 - v1))))
        # and the oracle's sweep count and bits, as cw_field_relax has
        want_sweeps = ref.relax(1e-10, 200_000)
        assert s1 == want_sweeps
        assert bits_equal(v1, ref.values)


def test_noalloc_relax_reports_the_budget_like_relax():
    ab = _abi()
    lib = ab.lib()
    ref = bathed_disc_inputs(40, 10, 0.004)
    g = gpu_from_oracle(ref)
    n = C.c_uint64()
    rc = lib.cw_field_relax_noalloc(g._h, 1e-12, 3, C.byref(n))
    assert rc == ab.CW_ERR_NOT_CONVERGED
    assert "did not settle within 3 sweeps" in ab.last_error()


def test_noalloc_relax_refuses_to_allocate_the_general_form():
    ab = _abi()
    lib = ab.lib()
    from cellweave_amd.field import Field

    g = Field(16, 16, 1, 0.2, 0.0, 1.0)
    g.set_source(1e-3)  # production and uptake kept apart: needs a buffer
    n = C.c_uint64()
    rc = lib.cw_field_relax_noalloc(g._h, 1e-8, 1000, C.byref(n))
    assert rc == ab.CW_ERR_ARG
    assert "allocating call" in ab.last_error()
    # an allocating call (here the upload that derives the coefficients)
    # creates it; then the allocation-free relax runs
    g2 = Field(16, 16, 1, 0.2, 0.0, 1.0)
    g2.set_source(1e-3)
    held = np.zeros(256, np.uint8)
    held[:16] = 1
    g.upload(held=held)
    g2.upload(held=held)
    s_alloc = g2.relax(1e-10, 100_000)
    assert lib.cw_field_relax_noalloc(g._h, 1e-10, 100_000, C.byref(n)) == ab.CW_OK
    assert int(n.value) == s_alloc
    assert bits_equal(g.download(), g2.download())
