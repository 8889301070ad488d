"""The reference's own ScalarField and coupling tests (crates/engine/src/pde.rs
tests module, crates/engine/src/coupling.rs tests module), run on the CPU
oracle. Names follow the reference's test functions."""
import math

import numpy as np
import pytest

from oracle.orf import ConfigError, OracleField, RelaxError

from helpers import D, bathed_disc_inputs, outside_disc


def test_a_pocket_walled_in_by_tissue_is_never_reached():  # pde.rs:833-845
    f = OracleField(11, 11, 1, 0.2, 0.0, 0.0)
    f.hold_bath(1.0, lambda x, y: not max(abs(x - 5), abs(y - 5)) == 3)
    assert f.concentration_at(5, 5) == 0.0
    assert f.concentration_at(0, 0) == 1.0


def test_one_gap_in_the_wall_lets_the_bath_in():  # pde.rs:848-858
    f = OracleField(11, 11, 1, 0.2, 0.0, 0.0)
    wall = lambda x, y: max(abs(x - 5), abs(y - 5)) == 3 and not (x == 5 and y == 2)
    f.hold_bath(1.0, lambda x, y: not wall(x, y))
    assert f.concentration_at(5, 5) == 1.0


def test_a_wall_that_only_touches_at_corners_still_seals():  # pde.rs:861-880
    f = OracleField(11, 11, 1, 0.2, 0.0, 0.0)
    f.hold_bath(1.0, lambda x, y: not abs(x - 5) + abs(y - 5) == 3)
    assert f.concentration_at(5, 5) == 0.0
    assert f.concentration_at(0, 0) == 1.0


def test_a_closed_box_keeps_what_it_holds():  # pde.rs:883-905
    f = OracleField(20, 20, 1, 0.2, 0.0, 0.0)
    v = f.values
    for y in range(2, 6):
        for x in range(2, 6):
            v[y * 20 + x] = 1.0
    before = f.values.sum()
    for _ in range(500):
        f.step(1.0)
    assert abs(f.values.sum() - before) < 1e-9


def test_an_unstable_time_step_is_refused():  # pde.rs:908-914
    with pytest.raises(ConfigError):
        OracleField(20, 20, 1, 0.3, 0.0, 1.0).step(1.0)


def test_a_decay_that_removes_more_than_everything_is_refused():  # pde.rs:917-923
    with pytest.raises(ConfigError):
        OracleField(20, 20, 1, 0.1, 2.0, 1.0).step(1.0)


def test_a_time_step_going_nowhere_is_refused():  # pde.rs:926-931
    f = OracleField(20, 20, 1, 0.1, 0.0, 1.0)
    with pytest.raises(ConfigError):
        f.step(0.0)
    with pytest.raises(ConfigError):
        f.step(-1.0)


def test_a_field_with_no_pixels_is_refused():  # pde.rs:934-938
    with pytest.raises(ConfigError):
        OracleField(0, 10, 1, 0.2, 0.0, 1.0)
    with pytest.raises(ConfigError):
        OracleField(10, 0, 1, 0.2, 0.0, 1.0)


def test_a_field_that_is_not_a_number_is_refused():  # pde.rs:941-949
    for diff, decay in ((float("nan"), 0.0), (0.2, float("nan")), (float("inf"), 0.0)):
        with pytest.raises(ConfigError):
            OracleField(10, 10, 1, diff, decay, 1.0)
    with pytest.raises(ConfigError):
        OracleField(10, 10, 1, 0.2, 0.0, 1.0).step(float("nan"))


def test_field_decays_toward_zero_without_source():  # pde.rs:952-961
    f = OracleField(20, 20, 1, 0.25, 0.1, 1.0)
    for _ in range(200):
        f.step(0.1)
    assert f.concentration_at(10, 10) < 0.5


def test_concentration_never_negative():  # pde.rs:964-979
    f = OracleField(30, 30, 1, 0.5, 2.0, 1.0)
    for _ in range(500):
        f.step(0.05)
    assert f.values.min() >= 0.0


def test_what_the_bath_hands_a_settled_tissue_is_what_it_takes_up():  # pde.rs:1003-1012
    f = bathed_disc_inputs(30, 6, 0.004)
    f.relax(1e-11, 50_000)
    taken = 0.004 * f.free_count()
    assert f.free_count() > 100
    assert abs(f.inflow() - taken) < 1e-6 * taken


def test_relaxing_settles_the_field_and_says_how_long_it_took():  # pde.rs:1015-1027
    f = bathed_disc_inputs(30, 6, 0.004)
    f.relax(1e-11, 50_000)
    settled = f.values.copy()
    assert f.relax(1e-11, 50_000) == 1
    assert np.max(np.abs(f.values - settled)) < 1e-6


def test_relaxing_refuses_to_pretend_when_it_cannot_settle_in_time():  # pde.rs:1030-1045
    f = bathed_disc_inputs(30, 6, 0.004)
    with pytest.raises(RelaxError):
        f.relax(1e-9, 3)


def test_a_held_pixel_is_a_boundary_condition_and_does_not_move():  # pde.rs:1048-1067
    f = bathed_disc_inputs(30, 6, 0.05)
    f.relax(1e-11, 50_000)
    for _ in range(200):
        f.step(1.0)
    assert f.concentration_at(0, 0) == 1.0
    assert f.concentration_at(15, 3) == 1.0
    assert f.concentration_at(15, 15) < 0.5


def test_clearing_the_uptake_lets_the_field_refill():  # pde.rs:1070-1081
    f = bathed_disc_inputs(30, 6, 0.01)
    f.relax(1e-11, 50_000)
    assert f.concentration_at(15, 15) < 0.9
    f.clear_uptake()
    f.relax(1e-7, 50_000)
    assert abs(f.concentration_at(15, 15) - 1.0) < 1e-5


def test_a_held_vessel_supplies_from_the_inside_out():  # pde.rs:1084-1113
    f = OracleField(60, 60, 1, D, 0.0, 0.0)
    vessel = lambda x, y: (x - 30) ** 2 + (y - 30) ** 2 <= 9
    for y in range(60):
        for x in range(60):
            if not vessel(x, y):
                f.set_uptake(x, y, 0.002)
    f.hold_pixels(1.0, vessel)
    f.relax(1e-7, 50_000)
    at = lambda r: f.concentration_at(30 + r, 30)
    assert at(0) == 1.0
    assert at(5) > at(10) > at(15)
    assert f.concentration_at(1, 1) < 0.01


def test_releasing_lets_a_former_bath_pixel_move_again():  # pde.rs:1116-1138
    f = OracleField(20, 20, 1, D, 0.0, 1.0)
    f.hold_bath(1.0, lambda x, y: True)
    for y in range(20):
        for x in range(20):
            f.set_uptake(x, y, 0.01)
    f.step(1.0)
    assert f.concentration_at(10, 10) == 1.0
    f.release_all()
    f.step(1.0)
    assert f.concentration_at(10, 10) < 1.0


def test_a_bath_and_a_vessel_can_hold_at_once():  # pde.rs:1141-1156
    f = OracleField(30, 30, 1, D, 0.0, 0.0)
    tissue = lambda x, y: 5 <= x < 25 and 5 <= y < 25
    f.hold_bath(1.0, lambda x, y: not tissue(x, y))
    f.hold_pixels(0.5, lambda x, y: x == 15 and y == 15)
    assert f.concentration_at(0, 0) == 1.0
    assert f.concentration_at(15, 15) == 0.5
    assert f.concentration_at(10, 10) == 0.0


def test_a_sweep_is_held_to_the_box_around_what_is_free():  # pde.rs:1159-1175
    f = OracleField(20, 20, 1, D, 0.0, 0.0)
    tissue = lambda x, y: 4 <= x < 9 and 11 <= y < 14
    f.hold_bath(1.0, lambda x, y: not tissue(x, y))
    assert f.free_span() == (4, 8, 11, 13, 0, 0)
    f.hold_pixels(1.0, lambda x, y: True)
    assert f.free_span() is None
    assert f.relax(1e-9, 10) == 1


def test_the_box_swept_leaves_what_lies_outside_it_as_it_was_held():  # pde.rs:1178-1200
    f = OracleField(40, 40, 1, D, 0.0, 1.0)
    disc = lambda x, y: (x - 20) ** 2 + (y - 20) ** 2 <= 64
    for y in range(40):
        for x in range(40):
            if disc(x, y):
                f.set_uptake(x, y, 0.001)
    f.hold_bath(1.0, lambda x, y: not disc(x, y))
    f.relax(1e-9, 100_000)
    for y in range(40):
        for x in range(40):
            if not disc(x, y):
                assert f.concentration_at(x, y) == 1.0
    assert f.concentration_at(20, 20) < 1.0


def test_a_bath_reaches_a_volume_through_its_six_faces():  # pde.rs:1232-1248
    column = (np.arange(512) % 64 == 4 * 8 + 4).astype(np.uint8)
    vol = OracleField(8, 8, 8, D, 0.0, 0.0)
    vol.hold_bath_at(1.0, column)
    assert vol.value_at(3 * 64 + 4 * 8 + 4) == 1.0
    assert vol.value_at(3 * 64 + 4 * 8 + 3) == 0.0
    flat = OracleField(8, 8, 1, D, 0.0, 0.0)
    flat.hold_bath_at(1.0, column[:64])
    assert flat.value_at(4 * 8 + 4) == 0.0


def test_setters_clamp_at_zero_and_ignore_out_of_range():  # pde.rs:146-182
    f = OracleField(5, 5, 1, D, 0.0, 0.0)
    f.set_uptake(1, 1, -3.0)
    f.set_production(2, 2, float("nan"))
    f.set_uptake(9, 9, 1.0)
    f.set_production_at(10_000, 1.0)
    assert f.uptake[6] == 0.0 and f.production[12] == 0.0
    assert f.uptake.sum() == 0.0 and f.production.sum() == 0.0


# ---- coupling.rs -----------------------------------------------------------

OBSTACLE = 0xFFFFFFFF


def cluster():
    """coupling.rs:211-223 in spirit: one cell in the middle of an 80x80
    lattice ringed by eight others, as discs of radius 6 (ids 1..9)."""
    grid = np.zeros((80, 80), np.uint32)
    yy, xx = np.mgrid[0:80, 0:80]
    centres = [(40, 40)] + [(int(round(40 + 13 * np.cos(k * np.pi / 4))),
                             int(round(40 + 13 * np.sin(k * np.pi / 4)))) for k in range(8)]
    for cid, (cx, cy) in enumerate(centres, start=1):
        grid[(xx - cx) ** 2 + (yy - cy) ** 2 <= 36] = cid
    return grid.reshape(-1), 10


def settled_field(grid, consumes):
    f = OracleField(80, 80, 1, 0.2, 0.0, 1.0)
    f.apply_uptake(grid, consumes, 0.003)
    f.hold_bath_at(1.0, (grid == 0).astype(np.uint8))
    f.relax(1e-7, 50_000)
    return f


def test_a_cell_in_the_middle_of_a_cluster_reads_less_oxygen_than_one_on_the_rim():
    grid, n = cluster()  # coupling.rs:234-254
    consumes = np.ones(n, np.uint8)
    consumes[0] = 0
    o2 = settled_field(grid, consumes).mean_per_cell(grid, n)
    inner, outer = o2[1], o2[2:].max()
    assert inner < outer
    assert inner < 0.9 and outer > inner + 0.05


def test_a_necrotic_cell_stops_consuming():  # coupling.rs:257-269
    grid, n = cluster()
    consumes = np.ones(n, np.uint8)
    consumes[0] = 0
    starved = settled_field(grid, consumes).mean_per_cell(grid, n)[1]
    consumes[1] = 0
    relieved = settled_field(grid, consumes).mean_per_cell(grid, n)[1]
    assert relieved > starved


def test_the_medium_reads_zero_and_every_cell_reads_a_mean():  # coupling.rs:272-279
    grid, n = cluster()
    consumes = np.ones(n, np.uint8)
    o2 = settled_field(grid, consumes).mean_per_cell(grid, n)
    assert o2.size == n and o2[0] == 0.0
    assert np.all((o2[1:] > 0) & (o2[1:] <= 1.0))


def test_gathers_skip_medium_obstacles_and_unknown_ids():  # coupling.rs:26-73
    grid = np.array([0, 1, 2, OBSTACLE, 7, 2], np.uint32)
    f = OracleField(6, 1, 1, 0.2, 0.0, 0.0)
    consumes = np.array([0, 1, 0], np.uint8)
    f.apply_uptake(grid, consumes, 0.5)
    assert list(f.uptake) == [0, 0.5, 0, 0, 0, 0]
    f.apply_background_uptake(grid, 0.25)
    assert list(f.uptake) == [0.25, 0.5, 0, 0, 0, 0]
    f.apply_production(grid, np.array([9.0, -1.0, 3.0]))
    assert list(f.production) == [0, 0, 3.0, 0, 0, 3.0]
    f.apply_uptake_per_cell(grid, np.array([1, 1, 1], np.uint8), np.array([0.0, 0.1, 0.2]))
    assert list(f.uptake) == [0, 0.1, 0.2, 0, 0, 0.2]


def _random_lattice(seed, w=41, h=29, ncells=37):
    """Ids over 0..ncells+4 (some past the table), obstacles, and medium."""
    rng = np.random.default_rng(seed)
    grid = rng.integers(0, ncells + 5, w * h).astype(np.uint32)
    grid[rng.random(w * h) < 0.1] = OBSTACLE
    grid[rng.random(w * h) < 0.15] = 0
    return grid, rng


def test_mean_per_cell_against_an_independent_sum():  # coupling.rs:164-182
    """Per id below the table: the sum of the values of its voxels, in
    increasing voxel order, over their count; zero for the medium, for ids
    with no voxel, and nothing for obstacles or ids past the table. A
    plausible slip (count + 1, the medium or obstacles counted, ids past the
    table wrapped) moves at least one entry."""
    for seed in range(4):
        grid, rng = _random_lattice(seed)
        ncells = 37
        f = OracleField(41, 29, 1, 0.2, 0.0, 0.0)
        f.values[:] = rng.uniform(0.0, 2.0, grid.size)
        got = f.mean_per_cell(grid, ncells)
        want = np.zeros(ncells)
        for k in range(1, ncells):
            idx = np.nonzero(grid == k)[0]
            s = 0.0
            for i in idx:  # voxel order, as the reference adds them
                s = s + float(f.values[i])
            want[k] = s / idx.size if idx.size else 0.0
        assert np.array_equal(got, want)
        # and against numpy's own reduction, to rounding
        ok = (grid > 0) & (grid < ncells)
        cnt = np.bincount(grid[ok], minlength=ncells)
        tot = np.bincount(grid[ok], weights=f.values[ok], minlength=ncells)
        ref = np.divide(tot, cnt, out=np.zeros(ncells), where=cnt > 0)
        ref[0] = 0.0
        assert np.allclose(got, ref, rtol=1e-14, atol=0.0)


def test_per_cell_gathers_against_numpy():  # coupling.rs:26-73
    """apply_uptake_per_cell and apply_production against a numpy statement:
    a voxel of a cell id below the table takes its cell's rate where that is
    positive (and, for uptake, where the cell's kind consumes); the medium,
    obstacles and ids past the table take nothing."""
    for seed in range(4):
        grid, rng = _random_lattice(100 + seed)
        ncells = 37
        rates = rng.uniform(-1e-3, 2e-3, ncells)
        rates[3] = float("nan")
        consumes = (rng.random(ncells) < 0.7).astype(np.uint8)
        consumes[0] = 0
        f = OracleField(41, 29, 1, 0.2, 0.0, 0.0)
        f.apply_uptake_per_cell(grid, consumes, rates)
        ok = (grid > 0) & (grid < ncells)
        ids = np.where(ok, grid, 0).astype(np.int64)
        r = np.where(ok, rates[ids], 0.0)
        want_up = np.where(ok & (consumes[ids] != 0) & (r > 0), r, 0.0)
        assert np.array_equal(f.uptake, want_up)
        f.apply_production(grid, rates)
        want_pr = np.where(ok & (r > 0), r, 0.0)
        assert np.array_equal(f.production, want_pr)


def still_medium_level(field, diffusion):
    """simulation.rs:2114-2125 settle_bath_level: 1 - inflow / (4 pi D R), R
    the radius of a ball of the free voxels' volume, floored at 0; 1 when
    nothing is free."""
    radius = (3.0 * field.free_count() / (4.0 * math.pi)) ** (1.0 / 3.0)
    if not radius > 0.0:
        return 1.0
    return max(1.0 - field.inflow() / (4.0 * math.pi * diffusion * radius), 0.0)


def test_a_still_medium_sits_below_the_source_by_what_the_tissue_takes_up():
    """simulation.rs:2657-2689: the first cell of the volume scenario
    (simulation.rs:2619-2637), a ball of 123 voxels at the centre of a 40^3
    box each taking up 0.01 in a medium held as a bath at 1, settled as
    Simulation::start settles it (1e-8); the level is the closed form within
    1e-6, and a medium oxygen crosses faster is depleted less."""
    n = 40
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:n]
    ball = (((xx - 20) ** 2 + (yy - 20) ** 2 + (zz - 20) ** 2) <= 9).reshape(-1)
    assert ball.sum() == 123
    f = OracleField(n, n, n, 0.2, 0.0, 1.0)
    f.uptake[:] = np.where(ball, 0.01, 0.0)
    f.hold_bath_at(1.0, (~ball).astype(np.uint8))
    f.relax(1e-8, 2_000_000, threads=4)
    radius = (3.0 * 123.0 / (4.0 * math.pi)) ** (1.0 / 3.0)
    expected = 1.0 - 0.01 * 123.0 / (4.0 * math.pi * 0.2 * radius)
    level = still_medium_level(f, 0.2)
    assert abs(level - expected) < 1e-6, (level, expected)
    faster = still_medium_level(f, 2.0)
    assert level < faster < 1.0
