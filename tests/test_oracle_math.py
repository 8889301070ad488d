"""Pinning the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors (its CI forbids data files), so the
oracle (a C restatement of crates/engine/src/pde.rs) is pinned here by
checks that do not retype its formula:
  - an assembled dense matrix solved by numpy (brute force),
  - a tridiagonal solve for W x 1 grids (scipy.linalg.solve_banded),
  - the uniform decay solution, the textbook disc and sphere profiles,
  - the complementarity conditions of the floored problem,
  - central finite differences for the adjoint gradients.
A plausible mistake in the oracle (a dropped neighbour, a wrong denominator,
a sign) fails the dense or tridiagonal comparisons.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle.orf import OracleField, optimal_omega

from helpers import D, bathed_disc_inputs


def assemble(w, h, d, diff, decay, source, held, values, uptake, production):
    """A c = b on the free voxels, from the PDE's definition: each free voxel
    balances diffusion to its real neighbours against decay and its sinks and
    sources; held voxels enter as data."""
    n = w * h * d
    free = np.flatnonzero(held == 0)
    pos = -np.ones(n, dtype=np.int64)
    pos[free] = np.arange(free.size)
    A = np.zeros((free.size, free.size))
    b = np.zeros(free.size)
    for k, i in enumerate(free):
        x, y, z = i % w, (i // w) % h, i // (w * h)
        nbrs = []
        if x > 0: nbrs.append(i - 1)
        if x < w - 1: nbrs.append(i + 1)
        if y > 0: nbrs.append(i - w)
        if y < h - 1: nbrs.append(i + w)
        if z > 0: nbrs.append(i - w * h)
        if z < d - 1: nbrs.append(i + w * h)
        A[k, k] = len(nbrs) * diff + decay
        b[k] = source + production[i] - uptake[i]
        for j in nbrs:
            if held[j]:
                b[k] += diff * values[j]
            else:
                A[k, pos[j]] -= diff
    return A, b, free


def random_case(rng, w, h, d, diff=0.3, decay=0.0, source=0.0, hold_frac=0.15):
    f = OracleField(w, h, d, diff, decay, 0.5)
    held = (rng.random(w * h * d) < hold_frac).astype(np.uint8)
    held[0] = 1  # at least one Dirichlet voxel
    vals = rng.uniform(0.5, 1.5, w * h * d)
    f.values[:] = vals
    f.held[:] = held
    # small sinks and sources: the floor stays inactive
    f.uptake[:] = rng.uniform(0, 2e-3, w * h * d)
    f.production[:] = np.where(rng.random(w * h * d) < 0.3, rng.uniform(0, 1e-3, w * h * d), 0.0)
    if source:
        f.set_source(source)
    f.copy_next_from_values()
    return f


@pytest.mark.parametrize("shape,decay,source", [((12, 9, 1), 0.0, 0.0), ((10, 7, 1), 0.05, 1e-3),
                                                ((6, 5, 4), 0.0, 0.0), ((5, 5, 5), 0.02, 0.0)])
def test_converged_sor_matches_dense_solve(shape, decay, source):
    rng = np.random.default_rng(sum(shape))
    w, h, d = shape
    f = random_case(rng, w, h, d, decay=decay, source=source)
    held = f.held.copy()
    vals0 = f.values.copy()
    A, b, free = assemble(w, h, d, f.diff, decay, source, held, vals0, f.uptake.copy(),
                          f.production.copy())
    want = np.linalg.solve(A, b)
    assert want.min() > 0, "case must keep the floor inactive"
    f.relax_residual(1e-15, 200_000, omega=1.5)
    got = f.values[free]
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # held voxels keep their value exactly
    assert np.array_equal(f.values[held == 1], vals0[held == 1])


def test_residual_vector_is_b_minus_Ac():
    rng = np.random.default_rng(3)
    f = random_case(rng, 9, 8, 3, decay=0.01)
    A, b, free = assemble(9, 8, 3, f.diff, 0.01, 0.0, f.held.copy(), f.values.copy(),
                          f.uptake.copy(), f.production.copy())
    r = f.residual_vector()[free]
    want = b - A @ f.values[free]
    assert np.allclose(r, want, rtol=1e-13, atol=1e-15)
    rel, r2, b2 = f.rel_residual()
    assert abs(rel - np.linalg.norm(want) / np.linalg.norm(b)) <= 1e-12 * rel


def test_a_row_matches_the_tridiagonal_solve():
    w = 40
    f = OracleField(w, 1, 1, 0.7, 0.01, 0.0)
    rng = np.random.default_rng(5)
    up = rng.uniform(0, 1e-4, w)
    f.uptake[:] = up
    mask = np.zeros(w, np.uint8)
    mask[0] = 1
    f.hold_where(2.0, mask)
    f.relax_residual(1e-15, 500_000, omega=1.8)
    # unknowns x_1..x_{w-1}; x_0 = 2 held; the last voxel has one real neighbour
    n = w - 1
    diag = np.full(n, 2 * 0.7 + 0.01)
    diag[-1] = 0.7 + 0.01
    ab = np.zeros((3, n))
    ab[0, 1:] = -0.7
    ab[1] = diag
    ab[2, :-1] = -0.7
    rhs = -up[1:].copy()
    rhs[0] += 0.7 * 2.0
    want = scipy.linalg.solve_banded((1, 1), ab, rhs)
    assert want.min() > 0, "case must keep the floor inactive"
    assert np.max(np.abs(f.values[1:] - want)) < 1e-11


def test_uniform_decay_without_holds():
    f = OracleField(11, 7, 3, 0.4, 0.2, 0.0)
    f.set_source(0.05)
    f.production[:] = 0.01
    f.uptake[:] = 0.02
    f.relax_residual(1e-14, 100_000, omega=1.2)
    assert np.allclose(f.values, (0.05 + 0.01 - 0.02) / 0.2, rtol=1e-12)


def test_bathed_disc_textbook_profile():
    """pde.rs:810-830: 1 - qR^2/(4D) within 12% of the drop."""
    radius, uptake = 8, 0.003
    f = bathed_disc_inputs(32, radius, uptake)
    f.relax(1e-11, 50_000)
    measured = f.concentration_at(16, 16)
    depth = uptake * radius * radius / (4 * D)
    assert abs(measured - (1 - depth)) / depth < 0.12


def test_box_size_does_not_change_the_tissue():
    """pde.rs:783-807: the same disc in a 30 and a 60 box agrees to 1e-9."""
    small = bathed_disc_inputs(30, 6, 0.004)
    small.relax(1e-11, 50_000)
    large = bathed_disc_inputs(60, 6, 0.004)
    large.relax(1e-11, 50_000)
    a, b = small.concentration_at(15, 15), large.concentration_at(30, 30)
    assert abs(a - b) < 1e-9
    assert 0.1 < a < 0.9


def test_starving_ball_sphere_formula():
    """pde.rs:1203-1229: 1 - uR^2/(6D) within 0.05."""
    side, r, u = 31, 10.0, 0.006
    c = side // 2
    zz, yy, xx = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
    inside = ((xx - c) ** 2 + (yy - c) ** 2 + (zz - c) ** 2 <= r * r).reshape(-1)
    f = OracleField(side, side, side, D, 0.0, 1.0)
    f.uptake[:] = np.where(inside, u, 0.0)
    f.hold_bath_at(1.0, (~inside).astype(np.uint8))
    f.relax(1e-10, 1_000_000)
    centre = f.value_at(c * (side * side + side + 1))
    assert abs(centre - (1 - u * r * r / (6 * D))) < 0.05


def test_floored_solution_satisfies_complementarity():
    """A disc starving hard enough that its core floors at zero: the oracle's
    fixed point is the obstacle problem's solution, c >= 0, w = A c - b >= 0,
    c * w = 0 on the free voxels."""
    f = bathed_disc_inputs(40, 15, 0.02)
    f.relax_residual(1e-13, 400_000, omega=1.9)
    A, b, free = assemble(40, 40, 1, f.diff, 0.0, 0.0, f.held.copy(), f.values.copy(),
                          f.uptake.copy(), f.production.copy())
    c = f.values[free]
    wv = A @ c - b
    assert (c == 0).sum() > 20, "the core must floor"
    assert c.min() >= 0.0
    assert wv.min() > -1e-10
    assert np.max(np.abs(c * wv)) < 1e-10
    # and the projected residual is what the criterion reads
    assert f.rel_residual()[0] <= 1e-13


def test_thread_count_does_not_change_a_bit():
    """pde.rs:982-1000, with OpenMP over rows in chunks of 4."""
    outs = []
    for threads in (1, 3, 8):
        f = bathed_disc_inputs(61, 20, 0.0005)
        n = f.relax(1e-11, 50_000, threads=threads)
        outs.append((n, f.values.copy().view(np.uint64)))
    for n, v in outs[1:]:
        assert n == outs[0][0]
        assert np.array_equal(v, outs[0][1])


def test_reference_omega_rule():
    """pde.rs:267-277: a plane is capped at 1.95, a volume takes the optimum."""
    f = bathed_disc_inputs(200, 60, 0.001)
    assert f.reference_omega() == 1.95
    g = OracleField(20, 20, 20, D, 0.0, 1.0)
    mask = np.zeros(8000, np.uint8)
    mask[:400] = 1
    g.hold_where(1.0, mask)
    assert g.reference_omega() == optimal_omega(20)


def test_adjoint_matches_finite_differences():
    """The CPU adjoint's gradients against central differences (FP64 solves
    converged to 1e-14)."""
    w = h = 14
    dd = 3
    rng = np.random.default_rng(11)
    grid = rng.integers(1, 6, size=w * h * dd).astype(np.uint32)
    ring = np.zeros((dd, h, w), np.uint8)
    ring[:, 0, :] = ring[:, -1, :] = ring[:, :, 0] = ring[:, :, -1] = 1
    ring[0] = ring[-1] = 1 if dd > 2 else ring[0]
    ring = ring.reshape(-1)
    rates = np.array([0, 1e-3, 2e-3, 1.5e-3, 0.5e-3, 1.2e-3])
    target = rng.uniform(0.9, 1.0, w * h * dd)

    def solve(diff, g, rate_vec):
        f = OracleField(w, h, dd, diff, 0.0, 1.0)
        f.uptake[:] = np.where(ring == 1, 0.0, rate_vec[grid])
        f.hold_where(g, ring)
        f.relax_residual(1e-15, 500_000, omega=1.6)
        return f

    def loss(f):
        return 0.5 * np.sum((f.values - target) ** 2)

    base = solve(0.3, 1.0, rates)
    assert base.values.min() > 0
    adj = base.adjoint(target, omega=1.6, rel_tol=1e-14, grid=grid, ncells=6)
    hD = 1e-6 * 0.3
    fd_D = (loss(solve(0.3 + hD, 1.0, rates)) - loss(solve(0.3 - hD, 1.0, rates))) / (2 * hD)
    assert abs(adj["dD"] - fd_D) <= 1e-6 * abs(fd_D)
    hg = 1e-6
    fd_g = (loss(solve(0.3, 1.0 + hg, rates)) - loss(solve(0.3, 1.0 - hg, rates))) / (2 * hg)
    assert abs(adj["dg"] - fd_g) <= 1e-6 * abs(fd_g)
    for k in (1, 3, 5):
        hk = 1e-6 * rates[k]
        up, dn = rates.copy(), rates.copy()
        up[k] += hk
        dn[k] -= hk
        fd = (loss(solve(0.3, 1.0, up)) - loss(solve(0.3, 1.0, dn))) / (2 * hk)
        assert abs(adj["du"][k] - fd) <= 1e-5 * abs(fd), (k, adj["du"][k], fd)


def test_views_keep_their_field_alive():
    """A view taken from a temporary field stays valid after the temporary
    is gone (the view holds the field)."""
    import gc

    def make():
        f = OracleField(32, 32, 1, 0.3, 0.0, 0.0)
        f.values[:] = np.arange(32 * 32, dtype=np.float64)
        return f

    v = make().values
    gc.collect()
    junk = [OracleField(32, 32, 1, 0.3, 0.0, 7.0) for _ in range(8)]
    for j in junk:
        j.values[:] = -1.0
    assert np.array_equal(v, np.arange(32 * 32, dtype=np.float64))
