"""Shapes at the edges of the tiling, and the full-size workloads.

Small and ragged shapes (one voxel, single rows and columns, widths either
side of the 64-element row pitch and of the 128-column tile, thin and tall
volumes) are compared with the oracle bit for bit under the reference's own
stopping rule. The full-size workloads, which the oracle cannot solve in
seconds, are checked through what does not depend on size: the oracle's FP64
relative residual of the GPU's answer (one pass), and the FP32+refine answer
against the FP64 one."""
import numpy as np
import pytest

from cellweave_amd import cases
from cellweave_amd.field import (CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64, Field,
                                 batch_relax)
from helpers import bits_equal, gpu_from_case, gpu_from_oracle, oracle_from_case
from oracle.orf import OracleField

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def ragged_case(w, h, d, seed):
    """Random sinks and sources, a held ring on the faces that exist, and a
    few held voxels inside."""
    rng = np.random.default_rng(seed)
    n = w * h * d
    ref = OracleField(w, h, d, 0.25, 0.01, 0.5)
    ref.uptake[:] = np.where(rng.random(n) < 0.5, rng.uniform(0, 2e-3, n), 0.0)
    ref.production[:] = np.where((rng.random(n) < 0.2) & (ref.uptake == 0),
                                 rng.uniform(0, 1e-3, n), 0.0)
    held = np.zeros((d, h, w), bool)
    if w > 2:
        held[:, :, 0] = held[:, :, -1] = True
    if h > 2:
        held[:, 0, :] = held[:, -1, :] = True
    held = held.reshape(-1) | (rng.random(n) < 0.03)
    ref.hold_where(0.9, held.astype(np.uint8))
    ref.copy_next_from_values()
    return ref


SHAPES = [(1, 1, 1), (1, 57, 1), (57, 1, 1), (63, 5, 1), (64, 5, 1), (65, 7, 1), (127, 9, 1),
          (128, 9, 1), (129, 11, 1), (257, 6, 1), (3, 3, 3), (5, 4, 2), (2, 3, 40), (70, 3, 2),
          (33, 31, 3)]


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
def test_ragged_shapes_bitwise(shape):
    say(f"case ragged {shape}")
    w, h, d = shape
    ref = ragged_case(w, h, d, sum(shape))
    g = gpu_from_oracle(ref)
    want = ref.relax(1e-11, 200_000)
    got = g.relax(1e-11, 200_000)
    assert got == want
    assert bits_equal(g.download(), ref.values)


@pytest.mark.parametrize("shape", [(65, 7, 1), (129, 11, 1), (5, 4, 2), (33, 31, 3)])
def test_ragged_shapes_refine(shape):
    say(f"case ragged refine {shape}")
    w, h, d = shape
    ref = ragged_case(w, h, d, 7 * sum(shape))
    g = gpu_from_oracle(ref)
    st = g.relax_ex(omega=1.7, criterion=CRIT_REL_RESIDUAL, tolerance=1e-9,
                    precision=PREC_FP32_REFINE, max_sweeps=200_000)
    chk = ragged_case(w, h, d, 7 * sum(shape))
    chk.values[:] = g.download()
    assert st["rel_residual"] <= 1e-9
    assert chk.rel_residual()[0] <= 1e-9


def residual_by_oracle(fc, values):
    chk = oracle_from_case(fc)
    chk.values[:] = values
    return chk.rel_residual()[0]


def test_full_size_planes_by_residual():
    """1024^2, three fields, at the bench's settings: the oracle's own FP64
    residual of each GPU answer, and FP32+refine against FP64."""
    say("case full-size 1024^2 three fields")
    fcs = cases.multifield_plane(n=1024)
    import math

    omega = 2.0 / (1.0 + math.sin(math.pi / 1024))
    g64 = [gpu_from_case(fc) for fc in fcs]
    g32 = [gpu_from_case(fc) for fc in fcs]
    s64, _ = batch_relax(g64, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=omega,
                         precision=PREC_FP64)
    s32, _ = batch_relax(g32, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=omega,
                         precision=PREC_FP32_REFINE)
    for fc, a, b, sa, sb in zip(fcs, g64, g32, s64, s32):
        va, vb = a.download(), b.download()
        say(f"  {fc.name}: fp64 sweeps {sa['sweeps']}, fp32 sweeps {sb['sweeps']}")
        assert residual_by_oracle(fc, va) <= 1e-8
        assert residual_by_oracle(fc, vb) <= 1e-8
        assert (va >= 0).all() and (vb >= 0).all()
        scale = max(float(np.max(np.abs(va))), 1e-30)
        assert np.max(np.abs(va - vb)) <= 1e-6 * scale


def test_full_size_volume_by_residual():
    """256^3 spheroid: the oracle's FP64 residual of the FP32+refine answer."""
    say("case full-size 256^3 spheroid")
    (fc,) = cases.spheroid()
    from cellweave_amd import campaign

    g = gpu_from_case(fc)
    st = g.relax_ex(omega=campaign.omega_for(fc), criterion=CRIT_REL_RESIDUAL, tolerance=1e-8,
                    precision=PREC_FP32_REFINE, max_sweeps=100_000)
    v = g.download()
    assert st["rel_residual"] <= 1e-8
    assert residual_by_oracle(fc, v) <= 1e-8
    assert (v >= 0).all()


def test_a_field_with_nothing_held_settles_on_decay():
    """No holds, no-flux walls everywhere: decay alone pins the level, the
    steady state of a uniform source s is s / decay in every voxel."""
    say("case nothing held")
    w, h = 97, 41
    g = Field(w, h, 1, 0.3, 0.05, 0.0)
    g.set_source(2e-3)
    st = g.relax_ex(omega=1.8, criterion=CRIT_REL_RESIDUAL, tolerance=1e-12,
                    max_sweeps=200_000)
    assert st["rel_residual"] <= 1e-12
    assert np.max(np.abs(g.download() - 2e-3 / 0.05)) <= 1e-9
