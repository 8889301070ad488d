"""Batched and graph-replayed solves: the same bits as field-by-field eager
solves, every replay checked against the eager result."""
import numpy as np
import pytest

from cellweave_amd import campaign, cases
from cellweave_amd.field import (CRIT_MAX_CHANGE, CRIT_REL_RESIDUAL, PREC_FP32_REFINE,
                                 PREC_FP64, batch_relax)
from helpers import bits_equal, gpu_from_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("crit,tol", [(CRIT_MAX_CHANGE, 1e-10), (CRIT_REL_RESIDUAL, 1e-9)])
def test_graph_replay_equals_eager(crit, tol):
    say(f"case graph vs eager crit={crit}")
    fcs = cases.multifield_plane(n=192) + cases.oxygen_disc_200()
    eager = [gpu_from_case(fc) for fc in fcs]
    st_e, _ = batch_relax(eager, criterion=crit, tolerance=tol, max_sweeps=200_000, omega=1.95)
    graph = [gpu_from_case(fc) for fc in fcs]
    # two solves in a row from the same start: the cached graph is replayed again
    for rep in range(2):
        for g, fc in zip(graph, fcs):
            g.upload(values=fc.initial_values())
        st_g, _ = batch_relax(graph, criterion=crit, tolerance=tol, max_sweeps=200_000,
                              omega=1.95, use_graph=1)
        for a, b, ge, gg in zip(st_e, st_g, eager, graph):
            assert a["sweeps"] == b["sweeps"], rep
            assert bits_equal(ge.download(), gg.download()), rep


def test_graph_refine_equals_eager():
    say("case graph refine")
    fcs = cases.multifield_plane(n=192)
    e = [gpu_from_case(fc) for fc in fcs]
    g = [gpu_from_case(fc) for fc in fcs]
    st_e, _ = batch_relax(e, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=1.96,
                          precision=PREC_FP32_REFINE)
    st_g, _ = batch_relax(g, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=1.96,
                          precision=PREC_FP32_REFINE, use_graph=1)
    for a, b, x, y in zip(st_e, st_g, e, g):
        assert a["sweeps"] == b["sweeps"]
        assert bits_equal(x.download(), y.download())


def test_fp64_planes_sweep_deep_launches():
    """FP64 planes take eight sweeps a launch by default (shape 9, 128 x 64
    regions) while the batch's tiles fit on the CUs at once, four (shape 8)
    beyond. The reference's stopping rule still stops at its own sweep (the
    verdict replays the tail), so the bits match the two-sweep shape's."""
    say("case fp64 depth 200^2 oxygen, max change")
    (fc,) = cases.oxygen_disc_200()
    g = gpu_from_case(fc)
    s = g.relax_ex(omega=1.9, criterion=CRIT_MAX_CHANGE, tolerance=1e-8, max_sweeps=200_000)
    assert s["tile_shape"] == 9 and s["sweeps_per_pass"] == 8
    v8 = g.download()
    g.fill(fc.initial)
    s2 = g.relax_ex(omega=1.9, criterion=CRIT_MAX_CHANGE, tolerance=1e-8, max_sweeps=200_000,
                    tile_shape=3)
    assert s2["sweeps_per_pass"] == 2 and s2["sweeps"] == s["sweeps"]
    assert bits_equal(g.download(), v8)
    g.fill(fc.initial)
    s4 = g.relax_ex(omega=1.9, criterion=CRIT_MAX_CHANGE, tolerance=1e-8, max_sweeps=200_000,
                    sweeps_per_pass=4)
    assert s4["tile_shape"] == 7 and s4["sweeps"] == s["sweeps"]
    assert bits_equal(g.download(), v8)
    say("case fp64 depth 3 x 1024^2 concurrent")
    fcs = cases.multifield_plane(n=1024)
    gs = [gpu_from_case(f) for f in fcs]
    st, _ = batch_relax(gs, criterion=CRIT_REL_RESIDUAL, tolerance=1e-3, omega=1.99)
    assert [x["tile_shape"] for x in st] == [8, 8, 8]
    assert [x["sweeps_per_pass"] for x in st] == [4, 4, 4]


@pytest.mark.parametrize("n,fp64_shape,fp32_shape",
                         [(64, 17, 17), (80, 17, 18), (96, 15, 18), (120, 9, 18), (160, 4, 18)])
def test_volume_chunk_follows_workgroups(n, fp64_shape, fp32_shape):
    """Volumes take the longest z chunk that still gives the batch enough
    workgroups: FP64 16 or 8 planes at 220, else 4, 2 or 1 plane at 150;
    FP32 spans of 2^18 voxels or more the two-sweep pipelined smoother with
    automatic chunks (shape 18), smaller ones 16 to 2 planes at 256, else 2
    or 1 plane at 150."""
    say(f"case volume chunk {n}^3")
    (fc,) = cases.spheroid(n=n, ncells=int(round(60000 * (n / 256) ** 3)), radius=110 * n / 256,
                           necrotic=40 * n / 256)
    g = gpu_from_case(fc)
    for prec, want in ((PREC_FP64, fp64_shape), (PREC_FP32_REFINE, fp32_shape)):
        g.fill(fc.initial)
        st, _ = batch_relax([g], criterion=CRIT_REL_RESIDUAL, tolerance=1e-3, omega=-1.0,
                            precision=prec)
        assert st[0]["tile_shape"] == want and st[0]["status"] == 0


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32_REFINE])
def test_small_campaign_converges_and_matches_single_solves(precision):
    say(f"case small campaign precision={precision}")
    hist = cases.campaign_histories(count=12, lo=48, hi=160)
    flat = [fc for h in hist for fc in h]
    fields = campaign.make_fields(flat)
    stats, secs = campaign.solve(fields, campaign.plan(flat, precision, table={}),
                                 precision=precision, tol=1e-8)
    for fc, st, f in zip(flat, stats, fields):
        assert st["status"] == 0
        assert st["rel_residual"] <= 1e-8
        assert st["omega"] == campaign.omega_for(fc)
    if precision == PREC_FP64:
        # the batch equals each field solved alone with its own factor
        for k in range(0, len(flat), 5):
            g = gpu_from_case(flat[k])
            st = g.relax_ex(criterion=CRIT_REL_RESIDUAL, tolerance=1e-8, omega=-1.0)
            assert st["omega"] == stats[k]["omega"]
            assert st["sweeps"] == stats[k]["sweeps"]
            assert bits_equal(g.download(), fields[k].download())


def test_plane_smoothing_depth_for_many_tiles():
    """With no sweep count asked for and more tiles than fit on the CUs at
    once, three concurrent 1024^2 fields take the depth with the least
    tile-work a sweep, tiles(S) x (20 + S) / S: 8 sweeps (363 tiles, against
    432, 507 and 768 at 10, 12 and 16). One free 2048^2 field, launched alone,
    takes the least time a sweep in whole rounds, rounds x (RY / 128) x
    (20 + S) / S: 12 sweeps on 128 x 160 regions (494 tiles, two rounds,
    6.67 against 7.0 for 8 sweeps on 128 x 128). MI355X has 256 CUs."""
    from cellweave_amd.field import Field

    say("case smoothing depth 3 x 1024^2")
    fcs = cases.multifield_plane(n=1024)
    gs = [gpu_from_case(fc) for fc in fcs]
    st, _ = batch_relax(gs, criterion=CRIT_REL_RESIDUAL, tolerance=1e-3, omega=1.99,
                        precision=PREC_FP32_REFINE)
    assert [s["sweeps_per_pass"] for s in st] == [8, 8, 8]
    assert [s["tile_shape"] for s in st] == [15, 15, 15]  # 128-row regions
    say("case smoothing depth 2048^2")
    g = Field(2048, 2048, 1, 0.3, 0.05, 0.0)
    g.set_source(2e-3)
    s = g.relax_ex(omega=1.0, criterion=CRIT_REL_RESIDUAL, tolerance=1e-6,
                   precision=PREC_FP32_REFINE, max_sweeps=100_000)
    assert s["sweeps_per_pass"] == 12 and s["tile_shape"] == 27
    assert np.max(np.abs(g.download() - 2e-3 / 0.05)) <= 1e-5 * (2e-3 / 0.05)


def test_default_fp32_depth_follows_the_batch():
    """With no sweep count asked for, an FP32 plane batch takes 64-row regions
    at 12 sweeps (shape 23) while they fit on the CUs twice over (two
    512-thread workgroups a CU), else 16 sweeps on 128-row regions (shape 24)
    while those fit at once, else the depth with the least tile-work a sweep
    (8 sweeps, shape 15, for three 1024^2 fields). Fields solved concurrently
    count together. The smoother's bits do not depend on the region shape."""
    say("case short regions 200^2 oxygen")
    (fc,) = cases.oxygen_disc_200()
    g = gpu_from_case(fc)
    s = g.relax_ex(omega=1.9, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8,
                   precision=PREC_FP32_REFINE)
    assert s["tile_shape"] == 23 and s["sweeps_per_pass"] == 12
    v_short = g.download()
    g.fill(fc.initial)
    s20 = g.relax_ex(omega=1.9, criterion=CRIT_REL_RESIDUAL, tolerance=1e-8,
                     precision=PREC_FP32_REFINE, tile_shape=20)
    assert s20["sweeps"] == s["sweeps"]
    assert np.array_equal(g.download(), v_short)
    for n, want_all, want_one in ((256, 23, 23), (512, 24, 23), (1024, 15, 24), (1536, 25, 15)):
        say(f"case default depth 3 x {n}^2 concurrent and one alone")
        fcs = cases.multifield_plane(n=n)
        gs = [gpu_from_case(fc) for fc in fcs]
        st, _ = batch_relax(gs, criterion=CRIT_REL_RESIDUAL, tolerance=1e-3, omega=1.9,
                            precision=PREC_FP32_REFINE)
        assert [x["tile_shape"] for x in st] == [want_all] * 3, (n, st)
        st1, _ = batch_relax(gs[:1], criterion=CRIT_REL_RESIDUAL, tolerance=1e-3, omega=1.9,
                             precision=PREC_FP32_REFINE)
        assert st1[0]["tile_shape"] == want_one, (n, st1)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_FP32_REFINE])
def test_a_few_fields_solve_concurrently_with_their_own_bits(precision):
    """Two to four fields run as concurrent per-field solves on their own
    streams: each gets the sweeps and bits it gets when solved alone with the
    tile shape the batch chose (the choice counts the whole batch's tiles)."""
    say(f"case concurrent batch precision={precision}")
    fcs = cases.multifield_plane(n=160)
    together = [gpu_from_case(fc) for fc in fcs]
    st_t, _ = batch_relax(together, criterion=CRIT_REL_RESIDUAL, tolerance=1e-9, omega=1.95,
                          precision=precision)
    for fc, g, s in zip(fcs, together, st_t):
        alone = gpu_from_case(fc)
        (s1,), _ = batch_relax([alone], criterion=CRIT_REL_RESIDUAL, tolerance=1e-9, omega=1.95,
                               precision=precision, tile_shape=s["tile_shape"])
        assert s["status"] == 0 and s1["status"] == 0
        assert s["sweeps"] == s1["sweeps"]
        assert bits_equal(g.download(), alone.download())


def test_a_larger_batch_keeps_one_launch_sequence_and_the_same_bits():
    say("case five-field batch")
    fcs = cases.multifield_plane(n=128) + cases.multifield_plane(seed=2100, n=96)[:2]
    together = [gpu_from_case(fc) for fc in fcs]
    st_t, _ = batch_relax(together, criterion=CRIT_MAX_CHANGE, tolerance=1e-10, omega=1.9)
    for fc, g, s in zip(fcs, together, st_t):
        alone = gpu_from_case(fc)
        n = alone.relax_ex(criterion=CRIT_MAX_CHANGE, tolerance=1e-10, omega=1.9)["sweeps"]
        assert s["sweeps"] == n
        assert bits_equal(g.download(), alone.download())


def test_a_batch_may_not_mix_planes_and_volumes():
    from cellweave_amd.field import Field

    say("case mixed batch")
    a = Field(16, 16, 1, 0.2, 0.01, 0.5)
    b = Field(8, 8, 8, 0.2, 0.01, 0.5)
    with pytest.raises(ValueError, match="mixes planes and volumes"):
        batch_relax([a, b], criterion=CRIT_REL_RESIDUAL, tolerance=1e-8)


def test_a_few_volumes_solve_concurrently_with_their_own_bits():
    say("case concurrent volumes")
    fcs = [cases.spheroid(seed=3100 + k, n=36 + 4 * k, ncells=40, radius=12.0 + k, necrotic=3.0)[0]
           for k in range(3)]
    together = [gpu_from_case(fc) for fc in fcs]
    st_t, _ = batch_relax(together, criterion=CRIT_MAX_CHANGE, tolerance=1e-10, omega=1.8)
    for fc, g, s in zip(fcs, together, st_t):
        alone = gpu_from_case(fc)
        n = alone.relax_ex(criterion=CRIT_MAX_CHANGE, tolerance=1e-10, omega=1.8)["sweeps"]
        assert s["sweeps"] == n
        assert bits_equal(g.download(), alone.download())


def test_a_concurrent_batch_reports_a_field_that_does_not_settle():
    from cellweave_amd.field import SimulationError

    say("case concurrent batch, one field capped")
    fcs = cases.multifield_plane(n=96)
    gs = [gpu_from_case(fc) for fc in fcs]
    with pytest.raises(SimulationError, match="did not settle"):
        batch_relax(gs, criterion=CRIT_REL_RESIDUAL, tolerance=1e-12, omega=1.9, max_sweeps=40)


@pytest.mark.parametrize("graph", [0, 1])
def test_a_batch_whose_fields_stop_at_different_times_keeps_their_own_bits(graph):
    """Eight fields of different sizes to a relative residual: as they stop
    the batch is laid out again over the rest, and every field still gets
    the sweeps and bits it gets alone at the tile shape the batch chose."""
    say(f"case uneven batch graph={graph}")
    hist = cases.campaign_histories(seed=4400, count=4, lo=48, hi=200)
    fcs = [fc for h in hist for fc in h]
    together = [gpu_from_case(fc) for fc in fcs]
    st_t, _ = batch_relax(together, criterion=CRIT_REL_RESIDUAL, tolerance=1e-9, omega=-1.0,
                          use_graph=graph)
    assert len({s["sweeps"] for s in st_t}) > 2
    for fc, g, s in zip(fcs, together, st_t):
        alone = gpu_from_case(fc)
        s1 = alone.relax_ex(criterion=CRIT_REL_RESIDUAL, tolerance=1e-9, omega=-1.0,
                            tile_shape=s["tile_shape"])
        assert s["sweeps"] == s1["sweeps"]
        assert bits_equal(g.download(), alone.download())
