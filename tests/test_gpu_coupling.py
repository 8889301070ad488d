"""GPU parity of the field's device-side state against the oracle: the gathers
from the cell-id lattice (coupling.rs:26-73), the holds and the bath flood
(pde.rs:465-631), free_count and inflow (pde.rs:415-461), the explicit step
(pde.rs:651-728) and mean_per_cell (coupling.rs:164-182). Integer and byte
results and every FP64 value the reference computes voxel by voxel are
compared bit for bit; inflow, a sum in another order, to a rounding bound."""
import numpy as np
import pytest

from cellweave_amd import cases
from cellweave_amd.field import ConfigError, Field, Lattice
from helpers import bits_equal
from oracle.orf import ConfigError as OracleConfigError
from oracle.orf import OracleField

pytestmark = pytest.mark.gpu

OBSTACLE = 0xFFFFFFFF


def say(msg):
    print(msg, flush=True)


def random_lattice(shape, ncells, seed, obstacles=0.02, medium=0.3):
    w, h, d = shape
    rng = np.random.default_rng(seed)
    n = w * h * d
    ids = rng.integers(1, ncells + 3, size=n).astype(np.uint32)  # a few ids past the table
    ids[rng.random(n) < medium] = 0
    ids[rng.random(n) < obstacles] = OBSTACLE
    return ids


def pair(shape, diff=0.3, decay=0.0, initial=1.0):
    w, h, d = shape
    return OracleField(w, h, d, diff, decay, initial), Field(w, h, d, diff, decay, initial)


def same_state(ref, g):
    held, up, pr = g.state()
    assert np.array_equal(held, ref.held.astype(np.uint8))
    assert bits_equal(up, ref.uptake)
    assert bits_equal(pr, ref.production)
    assert bits_equal(g.download(), ref.values)


SHAPES = [(96, 80, 1), (24, 20, 18)]


@pytest.mark.parametrize("shape", SHAPES)
def test_gathers_bitwise(shape):
    say(f"case gathers {shape}")
    ncells = 40
    ids = random_lattice(shape, ncells, 11)
    rng = np.random.default_rng(12)
    consumes = (rng.random(ncells) < 0.7).astype(np.uint8)
    consumes[0] = 0
    rates = rng.uniform(-1e-3, 3e-3, ncells)  # some negative: not taken up
    prod = rng.uniform(-1e-3, 2e-3, ncells)
    ref, g = pair(shape)
    lat = Lattice(*shape)
    lat.upload(ids)
    # stale sinks and sources that the gathers must clear
    ref.uptake[:] = 5.0
    ref.production[:] = 7.0
    g.upload(uptake=np.full(ref.n, 5.0), production=np.full(ref.n, 7.0))
    ref.apply_uptake(ids, consumes, 2.5e-3)
    g.apply_uptake(lat, consumes, 2.5e-3)
    ref.apply_production(ids, prod)
    g.apply_production(lat, prod)
    same_state(ref, g)
    ref.apply_background_uptake(ids, 4e-4)
    g.apply_background_uptake(lat, 4e-4)
    same_state(ref, g)
    ref.apply_uptake_per_cell(ids, consumes, rates)
    g.apply_uptake_per_cell(lat, rates, consumes)
    same_state(ref, g)
    # NaN and negative rates clamp to zero (pde.rs:170-182)
    ref.apply_uptake(ids, consumes, float("nan"))
    g.apply_uptake(lat, consumes, float("nan"))
    ref.apply_background_uptake(ids, -1.0)
    g.apply_background_uptake(lat, -1.0)
    same_state(ref, g)


@pytest.mark.parametrize("shape", SHAPES)
def test_holds_and_release_bitwise(shape):
    say(f"case holds {shape}")
    ids = random_lattice(shape, 30, 21, obstacles=0.05)
    ref, g = pair(shape)
    lat = Lattice(*shape)
    lat.upload(ids)
    ref.hold_where(0.7, ids == OBSTACLE)
    g.hold_obstacles_at(lat, 0.7)
    same_state(ref, g)
    mask = np.random.default_rng(22).random(ref.n) < 0.1
    ref.hold_where(0.3, mask)
    g.hold_where(mask, 0.3)
    same_state(ref, g)
    assert g.free_count() == ref.free_count()
    ref.release_all()
    g.release_all()
    same_state(ref, g)
    assert g.free_count() == ref.n


def flood_pair(shape, open_mask, value, prior=None):
    ref, g = pair(shape, initial=0.25)
    if prior is not None:
        ref.hold_where(0.9, prior)
        g.hold_where(prior, 0.9)
    ref.hold_bath_at(value, open_mask)
    g.hold_bath_at(open_mask, value)
    return ref, g


@pytest.mark.parametrize("shape,p_open", [((300, 200, 1), 0.62), ((300, 200, 1), 0.75),
                                          ((70, 40, 30), 0.72), ((70, 40, 30), 0.85)])
def test_bath_flood_percolating_masks_bitwise(shape, p_open):
    """Random masks near the percolation threshold make long winding paths
    across many flood tiles; the reachable set must equal the oracle's BFS."""
    say(f"case bath flood {shape} p_open={p_open}")
    rng = np.random.default_rng(int(p_open * 100) + shape[2])
    n = shape[0] * shape[1] * shape[2]
    open_mask = rng.random(n) < p_open
    ref, g = flood_pair(shape, open_mask, 1.0)
    same_state(ref, g)
    # with earlier holds, which the flood neither enters nor crosses
    prior = rng.random(n) < 0.05
    ref, g = flood_pair(shape, open_mask, 1.0, prior=prior)
    same_state(ref, g)


def walls():
    """The reference's bath fixtures (pde.rs:833-905, 1232-1248)."""
    out = []
    n = 20
    yy, xx = np.mgrid[0:n, 0:n]
    ring = (np.abs(xx - 10) == 5) & (np.abs(yy - 10) <= 5) | (np.abs(yy - 10) == 5) & (
        np.abs(xx - 10) <= 5)
    out.append(("pocket walled in", (n, n, 1), ~ring.reshape(-1)))
    gap = ring.copy()
    gap[5, 10] = False
    out.append(("one gap", (n, n, 1), ~gap.reshape(-1)))
    diamond = (np.abs(xx - 10) + np.abs(yy - 10) == 4)
    out.append(("corners seal", (n, n, 1), ~diamond.reshape(-1)))
    closed = np.zeros((n, n), bool)
    out.append(("nothing closed", (n, n, 1), ~closed.reshape(-1)))
    m = 12
    zz, yy3, xx3 = np.mgrid[0:m, 0:m, 0:m]
    shell = (np.maximum(np.maximum(np.abs(xx3 - 6), np.abs(yy3 - 6)), np.abs(zz - 6)) == 3)
    out.append(("closed shell in a volume", (m, m, m), ~shell.reshape(-1)))
    ball = ((xx3 - 6) ** 2 + (yy3 - 6) ** 2 + (zz - 6) ** 2) <= 9
    out.append(("ball in a volume", (m, m, m), ~ball.reshape(-1)))
    return out


@pytest.mark.parametrize("name,shape,open_mask", walls(), ids=[w[0] for w in walls()])
def test_bath_fixtures_bitwise(name, shape, open_mask):
    say(f"case bath fixture {name}")
    ref, g = flood_pair(shape, open_mask, 1.0)
    same_state(ref, g)


def test_medium_bath_from_the_lattice_bitwise():
    say("case hold_medium_bath multi-field lattice 512^2")
    (fc, *_rest) = cases.multifield_plane(n=512)
    ids = fc.grid.astype(np.uint32)
    ref, g = pair((512, 512, 1), initial=0.5)
    lat = Lattice(512, 512, 1)
    lat.upload(ids)
    ref.hold_bath_at(1.0, ids == 0)
    g.hold_medium_bath(lat, 1.0)
    same_state(ref, g)


@pytest.mark.parametrize("shape", SHAPES)
def test_inflow_and_free_count(shape):
    say(f"case inflow {shape}")
    rng = np.random.default_rng(31)
    ref, g = pair(shape, diff=0.37)
    vals = rng.uniform(0.0, 1.0, ref.n)
    held = (rng.random(ref.n) < 0.3).astype(np.uint8)
    ref.values[:] = vals
    ref.held[:] = held
    g.upload(values=vals, held=held)
    assert g.free_count() == ref.free_count()
    want = ref.inflow()
    got = g.inflow()
    # the same terms in another order: |error| <= (terms) * eps * |terms|
    w, h, d = shape
    v3 = vals.reshape(d, h, w)
    h3 = held.reshape(d, h, w).astype(bool)
    mag = 0.0
    for ax in range(3):
        if v3.shape[ax] < 2:
            continue
        a = [slice(None)] * 3
        b = [slice(None)] * 3
        a[ax] = slice(1, None)
        b[ax] = slice(None, -1)
        diff = np.abs(v3[tuple(a)] - v3[tuple(b)])
        cross = h3[tuple(a)] != h3[tuple(b)]
        mag += 0.37 * float(np.sum(diff[cross]))
    assert abs(got - want) <= 1e-12 * mag


def test_inflow_of_a_settled_tissue_is_its_uptake():
    """pde.rs:1003-1012: at a settled field the bath hands the tissue what it
    takes up."""
    say("case inflow settled disc")
    n, r, u = 60, 12, 0.004
    yy, xx = np.mgrid[0:n, 0:n]
    disc = ((xx - 30) ** 2 + (yy - 30) ** 2 <= r * r).reshape(-1)
    g = Field(n, n, 1, 0.2, 0.0, 1.0)
    g.upload(uptake=np.where(disc, u, 0.0))
    g.hold_bath_at(~disc, 1.0)
    g.relax(1e-13, 500_000)
    taken = u * disc.sum()
    assert abs(g.inflow() - taken) <= 1e-9 * taken


@pytest.mark.parametrize("shape", SHAPES)
def test_explicit_step_bitwise(shape):
    say(f"case step {shape}")
    rng = np.random.default_rng(41)
    ref, g = pair(shape, diff=0.15, decay=0.02)
    ref.set_source(1e-4)
    g.set_source(1e-4)
    vals = rng.uniform(0.0, 1.0, ref.n)
    held = (rng.random(ref.n) < 0.1).astype(np.uint8)
    up = np.where(rng.random(ref.n) < 0.5, rng.uniform(0, 0.3, ref.n), 0.0)
    pr = np.where(rng.random(ref.n) < 0.3, rng.uniform(0, 0.02, ref.n), 0.0)
    ref.values[:] = vals
    ref.held[:] = held
    ref.uptake[:] = up
    ref.production[:] = pr
    g.upload(values=vals, held=held, uptake=up, production=pr)
    dt = 0.9 * (0.25 * (2 / 3 if shape[2] > 1 else 1.0)) / 0.15
    for _ in range(7):
        ref.step(dt)
        g.step(dt)
    assert bits_equal(g.download(), ref.values)
    assert (g.download() >= 0).all()


def test_unstable_steps_are_refused():
    """pde.rs:908-931."""
    say("case step refused")
    g = Field(10, 10, 1, 1.0, 0.5, 1.0)
    ref = OracleField(10, 10, 1, 1.0, 0.5, 1.0)
    for dt in (0.3, 0.0, -1.0, float("nan")):
        with pytest.raises(ConfigError):
            g.step(dt)
        with pytest.raises(OracleConfigError):
            ref.step(dt)
    g3 = Field(6, 6, 6, 1.0, 0.0, 1.0)
    with pytest.raises(ConfigError):
        g3.step(0.2)  # above 0.25 * 2/3 in a volume
    g4 = Field(6, 6, 1, 0.1, 3.0, 1.0)
    with pytest.raises(ConfigError):
        g4.step(0.5)  # decay * dt above 1


@pytest.mark.parametrize("shape", SHAPES)
def test_mean_per_cell_bitwise(shape):
    say(f"case mean_per_cell {shape}")
    ncells = 50
    ids = random_lattice(shape, ncells, 51)
    ref, g = pair(shape)
    vals = np.random.default_rng(52).uniform(0.0, 1.0, ref.n)
    ref.values[:] = vals
    g.upload(values=vals)
    lat = Lattice(*shape)
    lat.upload(ids)
    for nc in (ncells, ncells + 3, 7):
        assert bits_equal(g.mean_per_cell(lat, nc), ref.mean_per_cell(ids, nc))
    # a new layout replaces the cached cell lists
    ids2 = random_lattice(shape, ncells, 53)
    lat.upload(ids2)
    assert bits_equal(g.mean_per_cell(lat, ncells), ref.mean_per_cell(ids2, ncells))


def test_lattice_size_mismatch_is_refused():
    say("case lattice mismatch")
    g = Field(10, 10, 1, 0.2, 0.0, 1.0)
    lat = Lattice(10, 11, 1)
    with pytest.raises(ValueError, match="does not match"):
        g.apply_background_uptake(lat, 1e-3)


def test_device_coupling_then_relax_equals_the_oracle():
    """The loop around the solve (simulation.rs:2138-2162 declare_sources, then
    relax, then mean_per_cell) done on the device, against the oracle doing
    the same on the host: values, sweep count and per-cell means bitwise."""
    say("case device coupling + relax, multi-field oxygen 256^2")
    oxygen, signal, _chem = cases.multifield_plane(n=256)
    ids = oxygen.grid.astype(np.uint32)
    ncells = int(ids[ids != OBSTACLE].max()) + 1
    consumes = np.ones(ncells, np.uint8)
    consumes[0] = 0
    consumes[np.random.default_rng(61).random(ncells) < 0.05] = 0  # necrotic
    prod = np.where(np.random.default_rng(62).random(ncells) < 0.3, 2e-4, 0.0)
    w = h = 256
    ref = OracleField(w, h, 1, oxygen.diff, 0.0, 1.0)
    g = Field(w, h, 1, oxygen.diff, 0.0, 1.0)
    lat = Lattice(w, h, 1)
    lat.upload(ids)
    ref.apply_uptake(ids, consumes, 1e-3)
    ref.release_all()
    ref.hold_bath_at(1.0, ids == 0)
    ref.copy_next_from_values()
    g.apply_uptake(lat, consumes, 1e-3)
    g.release_all()
    g.hold_medium_bath(lat, 1.0)
    same_state(ref, g)
    want = ref.relax(1e-9, 200_000)
    got = g.relax(1e-9, 200_000)
    assert got == want
    assert bits_equal(g.download(), ref.values)
    assert bits_equal(g.mean_per_cell(lat, ncells), ref.mean_per_cell(ids, ncells))
    # the signal field: production from the per-cell table, decay, no holds
    ref2 = OracleField(w, h, 1, signal.diff, signal.decay, 0.0)
    g2 = Field(w, h, 1, signal.diff, signal.decay, 0.0)
    ref2.apply_production(ids, prod)
    g2.apply_production(lat, prod)
    ring = np.zeros((h, w), np.uint8)
    ring[0, :] = ring[-1, :] = ring[:, 0] = ring[:, -1] = 1
    ref2.hold_where(0.0, ring.reshape(-1))
    g2.hold_where(ring.reshape(-1), 0.0)
    ref2.copy_next_from_values()
    want = ref2.relax(1e-12, 200_000)
    got = g2.relax(1e-12, 200_000)
    assert got == want
    assert bits_equal(g2.download(), ref2.values)


def test_changed_voxels_give_the_lattice_a_full_upload_gives():
    """cw_lattice_set with a step's changed voxels leaves the lattice as a
    full upload of the new grid does: the same gathers and means, bit for bit."""
    from cellweave_amd import workloads as wl

    print("case lattice set", flush=True)
    n = 40
    seq, ncells = wl.tissue_steps(n=n, radius=15.0, steps=2, copies=200, seed=6300)
    rates = np.random.default_rng(5).uniform(1e-4, 2e-3, ncells + 1)
    rates[0] = 0.0
    by_set, by_upload = Lattice(n, n, n), Lattice(n, n, n)
    by_set.upload(seq[0])
    for a, b in zip(seq, seq[1:]):
        changed = np.flatnonzero(a != b)
        by_set.set(changed, b[changed])
    by_upload.upload(seq[-1])
    f = Field(n, n, n, 20.0, 0.0, 1.0)
    g = Field(n, n, n, 20.0, 0.0, 1.0)
    f.apply_uptake_per_cell(by_set, rates)
    g.apply_uptake_per_cell(by_upload, rates)
    assert bits_equal(f.state()[1], g.state()[1])
    f.relax(1e-9, 200_000)
    assert bits_equal(f.mean_per_cell(by_set, ncells + 1), f.mean_per_cell(by_upload, ncells + 1))


def test_a_voxel_listed_twice_keeps_its_last_id():
    print("case lattice set twice", flush=True)
    lat = Lattice(8, 6, 1)
    lat.upload(np.ones(48, np.uint32))
    lat.set([5, 9, 5], [2, 3, 4])
    prod = np.zeros(5)
    prod[4] = 1e-3
    f = Field(8, 6, 1, 0.2, 0.0, 0.0)
    f.apply_production(lat, prod)
    want = np.zeros(48)
    want[5] = 1e-3
    assert bits_equal(f.state()[2], want)


def test_a_voxel_outside_the_lattice_is_refused():
    print("case lattice set out of range", flush=True)
    lat = Lattice(8, 6, 1)
    with pytest.raises(ValueError, match="outside"):
        lat.set([48], [1])
