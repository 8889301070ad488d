"""The committed golden fixtures (tests/golden/*.npz, made by
tests/golden/make_fixtures.py): their SHA-256 sums; that the seeded
generators still make the same inputs; that the oracle still gives the same
outputs; and the outputs against restatements that share no code with the
oracle:
  - numpy red-black sweeps in the reference's expression order
    (pde.rs:329-405: neighbours x-1, x+1, y-1, y+1, z-1, z+1, then
    ((d*sum + source) + production) - uptake over real*d + decay, the floor),
    run from the initial state, reproduce both stopping rules' answers bit for
    bit and stop at the same sweep: the largest change of the last sweep is
    below 1e-8 and of the one before it is not (pde.rs:279-283); the
    relative residual (numpy sums) is above 1e-8 one sweep before the stop
    and at or below it at the stop;
  - the per-cell gathers and means of the disc fixture against numpy
    (coupling.rs:26-73, 164-182);
  - the adjoint gradients against a sparse direct solve of A lambda = c - t.
"""
import hashlib
import importlib.util
import math
import os

import numpy as np
import pytest

from helpers import bits_equal

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
NAMES = ("disc64", "multifield64", "spheroid16", "campaign", "adjoint16")
TOL = 1e-8


def generator():
    spec = importlib.util.spec_from_file_location("make_fixtures",
                                                  os.path.join(GOLDEN, "make_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def fields(z):
    for k in range(int(z["count"])):
        p = f"f{k}_"
        yield k, {key[len(p):]: z[key] for key in z.files if key.startswith(p)}


def test_sha256_sums():
    with open(os.path.join(GOLDEN, "SHA256SUMS")) as fh:
        sums = dict(reversed(line.split()) for line in fh if line.strip())
    assert sorted(sums) == sorted(n + ".npz" for n in NAMES)
    for fname, want in sums.items():
        with open(os.path.join(GOLDEN, fname), "rb") as fh:
            assert hashlib.sha256(fh.read()).hexdigest() == want, fname


@pytest.mark.parametrize("name", NAMES)
def test_generators_make_the_same_inputs(name):
    mf = generator()
    fcs, rule = mf.fixture_cases()[name]
    z = load(name)
    assert int(z["count"]) == len(fcs)
    for (k, f), fc in zip(fields(z), fcs):
        assert tuple(f["dims"]) == (fc.w, fc.h, fc.d)
        assert str(f["name"]) == fc.name
        assert float(f["diff"]) == fc.diff and float(f["decay"]) == fc.decay
        assert np.array_equal(f["held"], fc.held)
        assert bits_equal(f["uptake"], fc.uptake) and bits_equal(f["production"], fc.production)
        assert bits_equal(f["values0"], fc.initial_values())
        assert float(f["omega"]) == mf.omega_of(fc, rule)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_gives_the_same_outputs(name):
    mf = generator()
    fresh = mf.build(name)
    z = load(name)
    assert sorted(fresh) == sorted(z.files)
    for key in z.files:
        assert bits_equal(np.asarray(fresh[key]).astype(z[key].dtype), z[key]) if \
            z[key].dtype.kind == "f" else np.array_equal(np.asarray(fresh[key]), z[key]), key


class Grid:
    """Neighbour views of a w x h x d grid in the reference's order."""

    def __init__(self, w, h, d):
        self.shape = (d, h, w)
        z, y, x = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
        self.colour = ((x + y + z) % 2).reshape(-1)
        self.dirs = []  # (valid mask, flat offset) for x-1, x+1, y-1, y+1, z-1, z+1
        for axis_coord, extent, step in ((x, w, 1), (y, h, w), (z, d, w * h)):
            self.dirs.append(((axis_coord > 0).reshape(-1), -step))
            self.dirs.append(((axis_coord < extent - 1).reshape(-1), step))
        self.real = sum(m.astype(np.float64) for m, _ in self.dirs)
        self.n = w * h * d
        self.idx = np.arange(self.n)

    def nbr(self, c, k):
        m, off = self.dirs[k]
        return m, c[np.where(m, self.idx + off, self.idx)]

    def nsum(self, c):
        s = np.zeros(self.n)
        for k in range(6):
            m, v = self.nbr(c, k)
            s = np.where(m, s + v, s)
        return s


def np_sweep(g, c, f, omega):
    """One red-black sweep: returns the new state and its largest change."""
    d, decay, src = float(f["diff"]), float(f["decay"]), float(f["source"])
    held = f["held"] != 0
    moved = 0.0
    for colour in (0, 1):
        s = g.nsum(c)
        bal = (((d * s + src) + f["production"]) - f["uptake"]) / (g.real * d + decay)
        after = (1.0 - omega) * c + omega * bal
        after = np.where(after > 0.0, after, 0.0)
        upd = (g.colour == colour) & ~held
        if upd.any():
            moved = max(moved, float(np.max(np.abs(after[upd] - c[upd]))))
        c = np.where(upd, after, c)
    return c, moved


def np_rel_residual(g, c, f):
    """||projected residual|| / ||b_eff|| over the free voxels (numpy sums)."""
    d, decay, src = float(f["diff"]), float(f["decay"]), float(f["source"])
    held = f["held"] != 0
    hs = np.zeros(g.n)
    for k in range(6):
        m, v = g.nbr(c, k)
        hs = np.where(m & held[np.where(m, g.idx + g.dirs[k][1], g.idx)], hs + v, hs)
    aii = g.real * d + decay
    b = (src + f["production"]) - f["uptake"]
    r = (b + d * g.nsum(c)) - aii * c
    rt = np.where((aii > 0) & ~(c + r / aii > 0.0), -aii * c, r)
    beff = b + d * hs
    free = ~held
    r2, b2 = float(np.sum(rt[free] ** 2)), float(np.sum(beff[free] ** 2))
    return math.sqrt(r2) / math.sqrt(b2) if b2 > 0 else math.sqrt(r2)


@pytest.mark.parametrize("name", NAMES)
def test_numpy_sweeps_reproduce_both_rules(name):
    z = load(name)
    for k, f in fields(z):
        w, h, d = (int(v) for v in f["dims"])
        g = Grid(w, h, d)
        omega = float(f["omega"])
        # the reference's rule: stop after the first sweep that moves less than tol
        c, moves = f["values0"].copy(), []
        for _ in range(int(f["maxchange_sweeps"])):
            c, m = np_sweep(g, c, f, omega)
            moves.append(m)
        assert moves[-1] < TOL and all(m >= TOL for m in moves[:-1]), (name, k)
        assert bits_equal(c, f["maxchange_values"]), (name, k)
        # the relative residual, tested before every sweep
        c = f["values0"].copy()
        n = int(f["residual_sweeps"])
        for _ in range(n - 1):
            c, _ = np_sweep(g, c, f, omega)
        assert np_rel_residual(g, c, f) > TOL, (name, k)
        c, _ = np_sweep(g, c, f, omega)
        rel = np_rel_residual(g, c, f)
        assert rel <= TOL and bits_equal(c, f["residual_values"]), (name, k)
        assert abs(rel - float(f["residual_rel"])) <= 1e-12 * rel, (name, k)


def test_disc_gathers_and_means_against_numpy():
    z = load("disc64")
    grid, ncells = z["grid"], int(z["ncells"])
    up, prod, consumes = z["cell_uptake"], z["cell_production"], z["consumes"]
    ok = (grid > 0) & (grid < ncells)
    ids = np.where(ok, grid, 0).astype(np.int64)
    want_up = np.where(ok & (consumes[ids] != 0) & (up[ids] > 0), up[ids], 0.0)
    assert np.array_equal(z["gathered_uptake"], want_up)
    want_pr = np.where(ok & (prod[ids] > 0), prod[ids], 0.0)
    assert np.array_equal(z["gathered_production"], want_pr)
    vals = z["f0_maxchange_values"]
    cnt = np.bincount(grid[ok], minlength=ncells)
    tot = np.bincount(grid[ok], weights=vals[ok], minlength=ncells)
    want = np.divide(tot, cnt, out=np.zeros(ncells), where=cnt > 0)
    want[0] = 0.0
    assert np.allclose(z["means"], want, rtol=1e-14, atol=0.0)
    assert (grid == 0xFFFFFFFF).sum() == 12 and (grid >= ncells).sum() == 24


def test_adjoint_gradients_against_a_direct_solve():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    z = load("adjoint16")
    (_, f), = fields(z)
    w, h, d = (int(v) for v in f["dims"])
    g = Grid(w, h, d)
    D, decay = float(f["diff"]), float(f["decay"])
    held = f["held"] != 0
    c, t, grid = z["solved"], z["target"], z["grid"]
    assert c[~held].min() > 0.0  # the floor is inactive: the gradients are the linear ones
    free = np.flatnonzero(~held)
    pos = -np.ones(g.n, np.int64)
    pos[free] = np.arange(free.size)
    rows, cols, vals = [free.copy()], [free.copy()], [(g.real * D + decay)[free]]
    for k in range(6):
        m, off = g.dirs[k]
        j = g.idx + off
        sel = free[m[free] & ~held[np.where(m, j, g.idx)][free]]
        rows.append(sel)
        cols.append(sel + off)
        vals.append(np.full(sel.size, -D))
    A = sp.csr_matrix((np.concatenate(vals), (pos[np.concatenate(rows)],
                                              pos[np.concatenate(cols)])),
                      shape=(free.size, free.size))
    lam = np.zeros(g.n)
    lam[free] = spl.spsolve(A.tocsc(), (c - t)[free])
    dsum = np.zeros(g.n)
    nheld = np.zeros(g.n)
    for k in range(6):
        m, v = g.nbr(c, k)
        dsum += np.where(m, v - c, 0.0)
        nheld += (m & held[np.where(m, g.idx + g.dirs[k][1], g.idx)]).astype(np.float64)
    dD = float(np.sum(lam[free] * dsum[free]))
    dg = float(D * np.sum(lam[free] * nheld[free]) + np.sum((c - t)[held]))
    ncells = int(z["ncells"])
    take = ~held & (grid > 0) & (grid < ncells) & (f["uptake"] > 0)
    du = -np.bincount(grid[take], weights=lam[take], minlength=ncells)
    assert abs(float(z["dl_dd"]) - dD) <= 1e-8 * abs(dD)
    assert abs(float(z["dl_dg"]) - dg) <= 1e-8 * abs(dg)
    assert np.max(np.abs(z["dl_du"] - du)) <= 1e-8 * np.max(np.abs(du))
