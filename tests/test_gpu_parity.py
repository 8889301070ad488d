"""GPU parity: the HIP solve against the CPU oracle through the C ABI.

The plain FP64 solve must equal the oracle bit for bit (values and sweep
count) on the reference's own fixtures (pde.rs tests) and on the synthetic
workloads; the mixed-precision solve must reach the FP64 residual target.
"""
import numpy as np
import pytest

from cellweave_amd import cases
from cellweave_amd.field import (CRIT_MAX_CHANGE, CRIT_REL_RESIDUAL, PREC_FP32_REFINE, PREC_FP64,
                                 Field, SimulationError, batch_relax)
from helpers import (D, bathed_disc_inputs, bits_equal, gpu_from_case, gpu_from_oracle,
                     oracle_from_case, outside_disc)
from oracle.orf import OracleField, RelaxError

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("size,radius,uptake,tol", [
    (30, 6, 0.004, 1e-11),     # pde.rs:783-807
    (32, 8, 0.003, 1e-11),     # pde.rs:810-830
    (61, 20, 0.0005, 1e-11),   # pde.rs:982-1000, odd box
    (60, 6, 0.004, 1e-11),
])
def test_bathed_disc_bitwise(size, radius, uptake, tol):
    say(f"case bathed_disc size={size} radius={radius} uptake={uptake}")
    ref = bathed_disc_inputs(size, radius, uptake)
    g = gpu_from_oracle(ref)
    want = ref.relax(tol, 50_000)
    got = g.relax(tol, 50_000)
    assert got == want
    assert bits_equal(g.download(), ref.values)


PLAIN_SHAPES = range(10)  # FP64 plain sweep kernels (LDS tiles and register-blocked)


@pytest.mark.parametrize("shape", PLAIN_SHAPES)
def test_oxygen_disc_200_max_change_bitwise(shape):
    say(f"case oxygen 200^2 omega=1.9 plain shape={shape}")
    (fc,) = cases.oxygen_disc_200()
    ref = oracle_from_case(fc)
    g = gpu_from_case(fc)
    want = ref.relax(1e-10, 100_000, omega=1.9)
    st = g.relax_ex(omega=1.9, criterion=CRIT_MAX_CHANGE, tolerance=1e-10, max_sweeps=100_000,
                    tile_shape=shape)
    assert st["sweeps"] == want
    assert bits_equal(g.download(), ref.values)


@pytest.mark.parametrize("shape", [3, 7, 8, 9])
@pytest.mark.parametrize("max_sweeps", [1, 2, 3, 5, 6, 7])
def test_budget_not_a_multiple_of_the_depth(shape, max_sweeps):
    """A run whose sweep budget is not a multiple of the launch depth ends
    with one-sweep launches: exactly max_sweeps sweeps run, as the oracle's."""
    say(f"case budget {max_sweeps} plain shape={shape}")
    (fc,) = cases.oxygen_disc_200()
    ref = oracle_from_case(fc)
    g = gpu_from_case(fc)
    with pytest.raises(RelaxError):
        ref.relax(1e-30, max_sweeps, omega=1.9)
    st, _ = batch_relax([g], check=False, omega=1.9, criterion=CRIT_MAX_CHANGE, tolerance=1e-30,
                        max_sweeps=max_sweeps, tile_shape=shape)
    assert st[0]["status"] != 0 and st[0]["sweeps"] == max_sweeps
    assert bits_equal(g.download(), ref.values)


@pytest.mark.parametrize("shape", [20, 23])
def test_fp32_budget_not_a_multiple_of_the_depth(shape):
    """FP32 smoothing capped by max_sweeps mid-segment runs exactly the cap."""
    say(f"case fp32 budget 37 shape={shape}")
    (fc,) = cases.oxygen_disc_200()
    g = gpu_from_case(fc)
    st, _ = batch_relax([g], check=False, omega=1.9, criterion=CRIT_REL_RESIDUAL, tolerance=1e-30,
                        max_sweeps=37, precision=PREC_FP32_REFINE, tile_shape=shape)
    assert st[0]["status"] != 0 and st[0]["sweeps"] == 37


def test_fp32_batch_budget_is_a_hard_limit():
    """A batch of five fields of different sizes (one launch sequence, the
    launches as long as the longest segment) under a max_sweeps that is not
    a multiple of the launch depth: no field runs past it."""
    say("case fp32 batch budget 101")
    hist = cases.campaign_histories(count=3, lo=96, hi=320, seed=4500)
    flat = [fc for h in hist for fc in h][:5]
    gs = [gpu_from_case(fc) for fc in flat]
    st, _ = batch_relax(gs, check=False, omega=-1.0, criterion=CRIT_REL_RESIDUAL, tolerance=1e-30,
                        max_sweeps=101, precision=PREC_FP32_REFINE)
    for s in st:
        assert s["status"] != 0 and s["sweeps"] <= 101, s
        assert s["sweeps"] > 101 - s["sweeps_per_pass"], s


@pytest.mark.parametrize("shape", PLAIN_SHAPES)
def test_general_form_every_shape(shape):
    say(f"case general form plain shape={shape}")
    ref, g = general_case()
    want = ref.relax(1e-12, 100_000)
    st = g.relax_ex(criterion=CRIT_MAX_CHANGE, tolerance=1e-12, max_sweeps=100_000,
                    tile_shape=shape)
    assert st["sweeps"] == want
    assert bits_equal(g.download(), ref.values)


def test_oxygen_disc_200_reference_dropin():
    say("case oxygen 200^2 reference relax (omega rule 1.95)")
    (fc,) = cases.oxygen_disc_200()
    ref = oracle_from_case(fc)
    g = gpu_from_case(fc)
    assert g.relax(1e-9, 100_000) == ref.relax(1e-9, 100_000)
    assert bits_equal(g.download(), ref.values)


@pytest.mark.parametrize("shape", [0, 2, 6])
def test_oxygen_disc_200_rel_residual_bitwise(shape):
    say(f"case oxygen 200^2 rel residual 1e-8 one sweep per pass, shape={shape}")
    (fc,) = cases.oxygen_disc_200()
    ref = oracle_from_case(fc)
    g = gpu_from_case(fc)
    want = ref.relax_residual(1e-8, 100_000, omega=1.9)
    st = g.relax_ex(omega=1.9, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, max_sweeps=100_000,
                    tile_shape=shape, check_every=1)
    assert st["sweeps"] == want
    assert bits_equal(g.download(), ref.values)
    rel = g.residual()
    assert rel <= 1e-8
    assert abs(rel - ref.rel_residual()[0]) <= 1e-12 * max(1.0, rel) + 1e-15


def general_case():
    """A uniform source and voxels that both take up and secrete: the
    unfolded ((d*sum + source) + p) - u form."""
    n = 40
    rng = np.random.default_rng(7)
    ref = OracleField(n, n, 1, 0.3, 0.01, 0.5)
    ref.set_source(2e-3)
    ref.uptake[:] = rng.uniform(0, 4e-3, n * n)
    ref.production[:] = np.where(rng.random(n * n) < 0.5, rng.uniform(0, 2e-3, n * n), 0.0)
    ring = np.zeros((n, n), np.uint8)
    ring[0, :] = ring[-1, :] = ring[:, 0] = ring[:, -1] = 1
    ref.hold_where(0.7, ring.reshape(-1))
    ref.copy_next_from_values()
    g = Field(n, n, 1, 0.3, 0.01, 0.5)
    g.set_source(2e-3)
    g.upload(values=ref.values.copy(), held=ref.held.copy(), uptake=ref.uptake.copy(),
             production=ref.production.copy())
    return ref, g


def test_starving_ball_3d_bitwise():
    """pde.rs:1203-1229: a ball taking up in a bath, in a volume."""
    say("case 3d ball 31^3")
    side, r, u = 31, 10.0, 0.006
    c = side // 2
    zz, yy, xx = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
    inside = ((xx - c) ** 2 + (yy - c) ** 2 + (zz - c) ** 2 <= r * r).reshape(-1)
    ref = OracleField(side, side, side, D, 0.0, 1.0)
    ref.uptake[:] = np.where(inside, u, 0.0)
    ref.hold_bath_at(1.0, (~inside).astype(np.uint8))
    ref.copy_next_from_values()
    g = gpu_from_oracle(ref)
    want = ref.relax(1e-10, 1_000_000)
    assert g.relax(1e-10, 1_000_000) == want
    assert bits_equal(g.download(), ref.values)
    centre = ref.value_at((side // 2) * (side * side + side + 1))
    assert abs(centre - (1.0 - u * r * r / (6.0 * D))) < 0.05


def test_max_sweeps_is_an_error():
    """pde.rs:1030-1045."""
    say("case max_sweeps 3")
    ref = bathed_disc_inputs(30, 6, 0.004)
    g = gpu_from_oracle(ref)
    with pytest.raises(SimulationError, match="did not settle within 3 sweeps"):
        g.relax(1e-9, 3)
    with pytest.raises(RelaxError):
        ref.relax(1e-9, 3)
    # the first three sweeps still ran, as the reference's loop runs them
    ref2 = bathed_disc_inputs(30, 6, 0.004)
    with pytest.raises(RelaxError):
        ref2.relax(1e-9, 3)
    assert bits_equal(g.download(), ref2.values)


def test_nothing_free_takes_one_sweep():
    """pde.rs:1159-1175: all held, relax is Ok(1)."""
    say("case nothing free")
    g = Field(20, 20, 1, D, 0.0, 0.0)
    g.upload(held=np.ones(400, np.uint8), values=np.ones(400))
    assert g.relax(1e-9, 10) == 1


def test_settled_field_takes_one_sweep():
    """pde.rs:1015-1027."""
    say("case settled re-relax")
    ref = bathed_disc_inputs(30, 6, 0.004)
    g = gpu_from_oracle(ref)
    g.relax(1e-11, 50_000)
    settled = g.download()
    assert g.relax(1e-11, 50_000) == 1
    assert np.max(np.abs(g.download() - settled)) < 1e-6


def test_batch_equals_single():
    say("case batch of three fields 256^2")
    fields = cases.multifield_plane(n=256)
    singles, batched = [], []
    for fc in fields:
        g = gpu_from_case(fc)
        st = g.relax_ex(criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, max_sweeps=200_000,
                        omega=1.97)
        singles.append((st["sweeps"], g.download()))
    gs = [gpu_from_case(fc) for fc in fields]
    stats, _ = batch_relax(gs, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, max_sweeps=200_000,
                           omega=1.97)
    for (sw, v), st, g in zip(singles, stats, gs):
        assert st["sweeps"] == sw
        assert bits_equal(g.download(), v)


SMOOTH_SHAPES = range(28)  # FP32 smoother kernels


@pytest.mark.parametrize("shape", SMOOTH_SHAPES)
def test_fp32_refine_reaches_residual_oxygen_200(shape):
    say(f"case oxygen 200^2 fp32 refine shape={shape}")
    (fc,) = cases.oxygen_disc_200()
    ref = oracle_from_case(fc)
    ref.relax_residual(1e-12, 200_000, omega=1.9)
    g = gpu_from_case(fc)
    st = g.relax_ex(omega=1.9, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, max_sweeps=200_000,
                    precision=PREC_FP32_REFINE, tile_shape=shape)
    say(f"  refine stats {st}")
    v = g.download()
    chk = oracle_from_case(fc)
    chk.values[:] = v
    rel = chk.rel_residual()[0]
    assert rel <= 1e-8, rel
    err = np.max(np.abs(v - ref.values))
    assert err < 1e-6, err
    assert st["outer_iters"] >= 1


SHAPES_3D = [*range(12), 15, 16, 17]  # in-place half sweeps (0) and z-streaming variants
SMOOTH_SHAPES_3D = range(19)  # and the two-sweep FP32 smoothers (12-14, 18)


def ball_case(side=31, r=10.0, u=0.006, general=False):
    c = side // 2
    zz, yy, xx = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
    inside = ((xx - c) ** 2 + (yy - c) ** 2 + (zz - c) ** 2 <= r * r).reshape(-1)
    ref = OracleField(side, side, side, D, 0.01 if general else 0.0, 1.0)
    ref.uptake[:] = np.where(inside, u, 0.0)
    if general:
        ref.set_source(1e-4)
        ref.production[:] = np.where(inside & (xx.reshape(-1) % 3 == 0), 2e-3, 0.0)
    ref.hold_bath_at(1.0, (~inside).astype(np.uint8))
    ref.copy_next_from_values()
    g = Field(side, side, side, D, ref.decay, 1.0)
    if general:
        g.set_source(1e-4)
    g.upload(values=ref.values.copy(), held=ref.held.copy(), uptake=ref.uptake.copy(),
             production=ref.production.copy())
    return ref, g


@pytest.mark.parametrize("shape", SHAPES_3D)
@pytest.mark.parametrize("general", [False, True])
def test_ball_3d_every_shape_bitwise(shape, general):
    say(f"case 3d ball shape={shape} general={general}")
    ref, g = ball_case(general=general)
    want = ref.relax(1e-10, 1_000_000)
    st = g.relax_ex(criterion=CRIT_MAX_CHANGE, tolerance=1e-10, max_sweeps=1_000_000,
                    tile_shape=shape)
    assert st["sweeps"] == want
    assert bits_equal(g.download(), ref.values)


@pytest.mark.parametrize("shape", [0, 1])
def test_box_3d_rel_residual(shape):
    """A 40x36x30 box, faces held, ragged against every tile size."""
    say(f"case 3d box rel residual shape={shape}")
    w, h, d = 40, 36, 30
    rng = np.random.default_rng(3)
    faces = np.zeros((d, h, w), np.uint8)
    faces[:, 0, :] = faces[:, -1, :] = faces[:, :, 0] = faces[:, :, -1] = 1
    faces[0] = faces[-1] = 1
    ref = OracleField(w, h, d, 0.25, 0.0, 1.0)
    ref.uptake[:] = np.where(faces.reshape(-1) == 1, 0.0, rng.uniform(0, 2e-3, w * h * d))
    ref.hold_where(1.0, faces.reshape(-1))
    ref.copy_next_from_values()
    g = gpu_from_oracle(ref)
    st = g.relax_ex(omega=1.85, criterion=CRIT_REL_RESIDUAL, tolerance=1e-10,
                    max_sweeps=100_000, tile_shape=shape, check_every=1)
    v = g.download()
    chk = OracleField(w, h, d, 0.25, 0.0, 1.0)
    chk.values[:] = v
    chk.uptake[:] = ref.uptake
    chk.held[:] = ref.held
    assert chk.rel_residual()[0] <= 1e-10
    # the same number of oracle sweeps gives the same bits
    for _ in range(st["sweeps"]):
        ref.sor_sweep(1.85)
    assert bits_equal(v, ref.values)


@pytest.mark.parametrize("shape", SMOOTH_SHAPES_3D)
def test_fp32_refine_3d(shape):
    say(f"case 3d refine shape={shape}")
    ref, g = ball_case(side=33, r=12.0)
    ref.relax_residual(1e-13, 1_000_000, omega=1.8)
    st = g.relax_ex(omega=1.8, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, max_sweeps=200_000,
                    precision=PREC_FP32_REFINE, tile_shape=shape)
    v = g.download()
    chk = OracleField(33, 33, 33, D, 0.0, 1.0)
    chk.values[:] = v
    chk.uptake[:] = ref.uptake
    chk.held[:] = ref.held
    assert chk.rel_residual()[0] <= 1e-8
    assert np.max(np.abs(v - ref.values)) < 1e-6


@pytest.mark.parametrize("inner_tol", [0.1, 0.03])
def test_fp32_refine_3d_in_place_odd_segments(inner_tol):
    """The in-place volume smoother keeps its answer in one buffer, so the
    correction must read it there whatever the parity of a segment's last
    pass; several segments of different lengths reach the oracle's answer
    (odd lengths: test_fp32_refine_3d_in_place_odd_cut)."""
    say(f"case 3d refine in place, inner_tol={inner_tol}")
    ref, g = ball_case(side=29, r=11.0)
    ref.relax_residual(1e-13, 1_000_000, omega=1.75)
    st = g.relax_ex(omega=1.75, criterion=CRIT_REL_RESIDUAL, tolerance=1e-10, max_sweeps=50_000,
                    precision=PREC_FP32_REFINE, tile_shape=0, inner_tolerance=inner_tol)
    say(f"  refine stats {st}")
    assert st["outer_iters"] >= 4
    assert np.max(np.abs(g.download() - ref.values)) < 1e-8


@pytest.mark.parametrize("max_sweeps", [33, 57])
def test_fp32_refine_3d_in_place_odd_cut(max_sweeps):
    """Volume segments are multiples of two sweeps, but max_sweeps can cut
    the last one to an odd length: the in-place smoother's correction still
    reads its answer from its one buffer. The ping-pong smoother, which
    stores each sweep in the other buffer, gives the same values to FP32
    rounding; a correction read from the wrong buffer would be off by a
    whole sweep's change."""
    say(f"case 3d refine in place, odd cut at {max_sweeps}")
    _, g0 = ball_case(side=29, r=11.0)
    _, g1 = ball_case(side=29, r=11.0)
    kw = dict(check=False, omega=1.75, criterion=CRIT_REL_RESIDUAL, tolerance=1e-14,
              max_sweeps=max_sweeps, precision=PREC_FP32_REFINE)
    s0, _ = batch_relax([g0], tile_shape=0, **kw)
    s1, _ = batch_relax([g1], tile_shape=1, **kw)
    assert s0[0]["sweeps"] == s1[0]["sweeps"] == max_sweeps
    assert s0[0]["status"] != 0 and s1[0]["status"] != 0
    v0, v1 = g0.download(), g1.download()
    moved = np.max(np.abs(v1 - 1.0))
    assert moved > 1e-3, moved
    assert np.max(np.abs(v0 - v1)) < 1e-6


@pytest.mark.parametrize("dims", [(90, 70, 1), (30, 26, 22)])
def test_repeated_relaxes_on_one_field_as_its_free_box_changes(dims):
    """cw_field_relax keeps its plan and captured graph between calls and
    captures again when the layout changes. Relax the same field eight times
    while the held region alternates between patterns with different free
    boxes (and so different tilings), each time from where the last solve
    left it, against the oracle doing the same: sweeps and bits every time."""
    w, h, d = dims
    say(f"case repeated relaxes {dims}")
    rng = np.random.default_rng(11 + d)
    n = w * h * d
    ref = OracleField(w, h, d, 0.25, 0.01, 0.5)
    ref.uptake[:] = np.where(rng.random(n) < 0.4, rng.uniform(0, 2e-3, n), 0.0)
    ref.copy_next_from_values()
    g = gpu_from_oracle(ref)
    zz, yy, xx = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    masks = []
    for k in range(4):
        m = (xx < 1 + 3 * k) | (yy >= h - 1 - 2 * k) | (xx == w - 1) | (yy == 0)
        if d > 1:
            m |= (zz == 0) | (zz >= d - 1 - k)
        masks.append(m.reshape(-1).astype(np.uint8))
    for step in range(8):
        m = masks[step % 4] if step % 2 == 0 else masks[3 - step % 4]
        level = 0.4 + 0.1 * step
        ref.release_all()
        ref.hold_where(level, m)
        g.release_all()
        g.hold_where(m, level)
        want = ref.relax(1e-9, 200_000)
        got = g.relax(1e-9, 200_000)
        say(f"  step {step}: sweeps gpu {got} oracle {want}")
        assert got == want
        assert bits_equal(g.download(), ref.values)


def test_two_sweep_volume_shapes_are_refused_for_fp64():
    say("case two-sweep shape in FP64")
    ref, _ = ball_case()
    g = gpu_from_oracle(ref)
    with pytest.raises(ValueError, match="FP32 smoothing only"):
        g.relax_ex(omega=1.8, criterion=CRIT_MAX_CHANGE, tolerance=1e-10, tile_shape=12)


@pytest.mark.parametrize("two", [12, 13, 14, 18])
def test_two_sweep_smoother_gives_the_one_sweep_bits(two):
    """Red-black SOR fixes every cell's operands, whatever order the stream
    visits planes in: the two-sweep FP32 volume smoother must give the bits
    of the one-sweep one (same 32-row region, same arithmetic), so the whole
    refinement matches, sweep for sweep."""
    say(f"case two-sweep shape {two} against one-sweep")
    ref, g1 = ball_case(side=37, r=14.0)
    _, g2 = ball_case(side=37, r=14.0)
    kw = dict(omega=1.85, criterion=CRIT_REL_RESIDUAL, tolerance=1e-9, max_sweeps=200_000,
              precision=PREC_FP32_REFINE)
    s1 = g1.relax_ex(tile_shape=1, **kw)
    s2 = g2.relax_ex(tile_shape=two, **kw)
    assert s2["sweeps_per_pass"] == 2
    assert s1["sweeps"] == s2["sweeps"]
    assert bits_equal(g1.download(), g2.download())


@pytest.mark.parametrize("two", [12, 18])
@pytest.mark.parametrize("dims", [(40, 36, 30), (131, 57, 23), (70, 90, 41)])
def test_two_sweep_smoother_on_ragged_boxes(two, dims):
    """Boxes held on their faces, so the free span starts on odd columns,
    rows and planes (the two-sweep kernel's regions start on even ones) and
    ends ragged against every region and chunk: the bits of the one-sweep
    smoother, sweep for sweep, with and without a lower bound."""
    w, h, d = dims
    say(f"case two-sweep shape {two} on a {dims} box")
    rng = np.random.default_rng(w + h + d)
    faces = np.zeros((d, h, w), np.uint8)
    faces[:, 0, :] = faces[:, -1, :] = faces[:, :, 0] = faces[:, :, -1] = 1
    faces[0] = faces[-1] = 1
    for sinks in (True, False):
        up = np.where(faces.reshape(-1) == 1, 0.0, rng.uniform(0, 2e-3, w * h * d)) if sinks \
            else np.zeros(w * h * d)
        gs = []
        for _ in range(2):
            ref = OracleField(w, h, d, 0.25, 0.0 if sinks else 1e-3, 1.0)
            ref.uptake[:] = up
            ref.hold_where(1.0, faces.reshape(-1))
            ref.copy_next_from_values()
            gs.append(gpu_from_oracle(ref))
        kw = dict(omega=1.9, criterion=CRIT_REL_RESIDUAL, tolerance=1e-9, max_sweeps=200_000,
                  precision=PREC_FP32_REFINE)
        s1 = gs[0].relax_ex(tile_shape=1, **kw)
        s2 = gs[1].relax_ex(tile_shape=two, **kw)
        assert s2["sweeps_per_pass"] == 2
        assert s1["sweeps"] == s2["sweeps"], (s1, s2)
        assert bits_equal(gs[0].download(), gs[1].download())


def test_fp32_volume_default_is_the_two_sweep_pipe():
    """FP32 volumes of 2^18 voxels or more take the pipelined two-sweep
    smoother with automatic chunks by default (smaller ones keep the
    one-sweep rule); the two-sweep default gives the one-sweep smoother's
    bits."""
    say("case default FP32 volume shape")
    for dims, want in (((72, 70, 66), 18), ((40, 36, 30), None)):
        w, h, d = dims
        faces = np.zeros((d, h, w), np.uint8)
        faces[:, 0, :] = faces[:, -1, :] = faces[:, :, 0] = faces[:, :, -1] = 1
        faces[0] = faces[-1] = 1
        up = np.where(faces.reshape(-1) == 1, 0.0, 1e-3)
        gs = []
        for _ in range(2):
            ref = OracleField(w, h, d, 0.25, 0.0, 1.0)
            ref.uptake[:] = up
            ref.hold_where(1.0, faces.reshape(-1))
            ref.copy_next_from_values()
            gs.append(gpu_from_oracle(ref))
        kw = dict(omega=1.9, criterion=CRIT_REL_RESIDUAL, tolerance=1e-9, max_sweeps=200_000,
                  precision=PREC_FP32_REFINE)
        s0 = gs[0].relax_ex(**kw)
        s1 = gs[1].relax_ex(tile_shape=1, **kw)
        assert s0["rel_residual"] <= 1e-9 and s1["rel_residual"] <= 1e-9
        if want is None:
            # the one-sweep rule (the in-place half sweeps round differently)
            assert s0["tile_shape"] != 18 and s0["sweeps_per_pass"] == 1
            continue
        assert s0["tile_shape"] == want and s0["sweeps_per_pass"] == 2
        assert s0["sweeps"] == s1["sweeps"]
        assert bits_equal(gs[0].download(), gs[1].download())


def test_fp32_volume_batch_gets_the_bits_alone():
    """Five volumes of different sizes in one batch (one launch sequence of
    two-sweep launches): each field gets the sweeps and bits it gets alone
    at the same shape, segments ending on launch boundaries; under a
    max_sweeps that is not a multiple of two, no field runs past it."""
    say("case fp32 volume batch of five")
    dims = [(40, 36, 30), (52, 44, 38), (66, 60, 58), (72, 70, 66), (48, 80, 70)]

    def make(w, h, d):
        faces = np.zeros((d, h, w), np.uint8)
        faces[:, 0, :] = faces[:, -1, :] = faces[:, :, 0] = faces[:, :, -1] = 1
        faces[0] = faces[-1] = 1
        rng = np.random.default_rng(w * h + d)
        ref = OracleField(w, h, d, 0.25, 0.0, 1.0)
        ref.uptake[:] = np.where(faces.reshape(-1) == 1, 0.0, rng.uniform(0, 2e-3, w * h * d))
        ref.hold_where(1.0, faces.reshape(-1))
        ref.copy_next_from_values()
        return gpu_from_oracle(ref)

    kw = dict(omega=-1.0, criterion=CRIT_REL_RESIDUAL, tolerance=1e-9, max_sweeps=200_000,
              precision=PREC_FP32_REFINE, tile_shape=18)
    batch = [make(*x) for x in dims]
    st, _ = batch_relax(batch, **kw)
    for x, s, g in zip(dims, st, batch):
        (one,) = [make(*x)]
        s1, _ = batch_relax([one], **kw)
        assert s["sweeps_per_pass"] == 2 and s["rel_residual"] <= 1e-9
        assert s["sweeps"] == s1[0]["sweeps"], (x, s, s1[0])
        assert bits_equal(g.download(), one.download()), x
    cap = [make(*x) for x in dims]
    st, _ = batch_relax(cap, check=False, **dict(kw, tolerance=1e-30, max_sweeps=101))
    for s in st:
        assert s["status"] != 0 and 100 <= s["sweeps"] <= 101, s
