"""Seeded synthetic workloads for the diffusion field solve.

Each builder returns FieldCase objects: a grid, a Cellular Potts style cell-id
lattice (u32 ids, 0 = medium, laid out x fastest as cpm.rs:194-198 stores it),
per-voxel uptake and production gathered from per-cell tables the way
coupling.rs:26-73 gathers them, and a Dirichlet hold on the outer layer of the
box (pde.rs:491-498, hold_where).

Units: one pixel is 10 um and one unit of time is 1 s, so a diffusion
coefficient D in um^2/s enters the solver as D / 100 px^2/s.

Seeds are s = 1000 * config + k. Voronoi cells: seeds uniform in a disc (ball);
every voxel of the disc (ball) takes its nearest seed, exact ties to the lower
id.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

DX_UM = 10.0


def d_px(d_um2: float) -> float:
    return d_um2 / (DX_UM * DX_UM)


@dataclasses.dataclass
class FieldCase:
    name: str
    w: int
    h: int
    d: int
    diff: float          # px^2 per s
    decay: float         # per s
    initial: float
    hold_value: float
    held: np.ndarray     # u8 [n]
    uptake: np.ndarray   # f64 [n]
    production: np.ndarray  # f64 [n]
    source: float = 0.0
    grid: np.ndarray | None = None          # u32 [n]
    per_cell_uptake: np.ndarray | None = None  # f64 [ncells]
    seed: int = 0

    @property
    def n(self) -> int:
        return self.w * self.h * self.d

    def initial_values(self) -> np.ndarray:
        v = np.full(self.n, self.initial, dtype=np.float64)
        v[self.held.astype(bool)] = self.hold_value
        return v

    def span(self):
        """Box around the free voxels (pde.rs:305-320), or None."""
        free = np.flatnonzero(self.held == 0)
        if free.size == 0:
            return None
        x = free % self.w
        y = (free // self.w) % self.h
        z = free // (self.w * self.h)
        return (int(x.min()), int(x.max()), int(y.min()), int(y.max()), int(z[0]), int(z[-1]))

    def span_voxels(self) -> int:
        s = self.span()
        if s is None:
            return 0
        return (s[1] - s[0] + 1) * (s[3] - s[2] + 1) * (s[5] - s[4] + 1)


def box_faces(w: int, h: int, d: int) -> np.ndarray:
    """Mask of the outer layer of the box: the ring of a plane, the faces of a volume."""
    m = np.zeros((d, h, w), dtype=np.uint8)
    m[:, 0, :] = 1
    m[:, -1, :] = 1
    m[:, :, 0] = 1
    m[:, :, -1] = 1
    if d > 1:
        m[0] = 1
        m[-1] = 1
    return m.reshape(-1)


def voronoi_disc(w, h, cx, cy, radius, nseeds, rng) -> np.ndarray:
    """u32 ids on a w x h plane: 0 outside the disc, 1..nseeds inside."""
    from scipy.spatial import cKDTree

    r = radius * np.sqrt(rng.random(nseeds))
    t = 2.0 * math.pi * rng.random(nseeds)
    seeds = np.stack([cx + r * np.cos(t), cy + r * np.sin(t)], axis=1)
    yy, xx = np.mgrid[0:h, 0:w]
    inside = (xx - cx) ** 2 + (yy - cy) ** 2 <= radius * radius
    pts = np.stack([xx[inside], yy[inside]], axis=1).astype(np.float64)
    ids = _nearest(cKDTree(seeds), pts)
    grid = np.zeros((h, w), dtype=np.uint32)
    grid[inside] = ids + 1
    return grid.reshape(-1)


def voronoi_ball(n, c, radius, nseeds, rng) -> np.ndarray:
    """u32 ids on an n^3 volume: 0 outside the ball, 1..nseeds inside."""
    from scipy.spatial import cKDTree

    v = rng.normal(size=(nseeds, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    r = radius * np.cbrt(rng.random(nseeds))
    seeds = c + v * r[:, None]
    tree = cKDTree(seeds)
    grid = np.zeros((n, n, n), dtype=np.uint32)
    yy, xx = np.mgrid[0:n, 0:n]
    for z in range(n):
        dz2 = (z - c) ** 2
        if dz2 > radius * radius:
            continue
        inside = (xx - c) ** 2 + (yy - c) ** 2 + dz2 <= radius * radius
        if not inside.any():
            continue
        pts = np.stack([xx[inside], yy[inside], np.full(inside.sum(), z)], axis=1).astype(np.float64)
        grid[z][inside] = _nearest(tree, pts) + 1
    return grid.reshape(-1)


def _nearest(tree, pts) -> np.ndarray:
    dist, idx = tree.query(pts, k=2, workers=-1)
    tie = dist[:, 0] == dist[:, 1]
    out = idx[:, 0].copy()
    out[tie] = np.minimum(idx[tie, 0], idx[tie, 1])
    return out


def gather(grid: np.ndarray, per_cell: np.ndarray) -> np.ndarray:
    """Per-voxel rate from a per-cell table, medium and unknown ids at zero
    (coupling.rs:61-73 shape, rates clamped at zero as pde.rs:170-182 does)."""
    table = np.concatenate([np.maximum(per_cell, 0.0), [0.0]])
    g = grid.astype(np.int64)
    g = np.where((g <= 0) | (g >= per_cell.size), per_cell.size, g)
    return table[g]


def _field(name, w, h, d, diff, decay, initial, hold, uptake, production, grid=None,
           per_cell_uptake=None, seed=0):
    return FieldCase(name=name, w=w, h=h, d=d, diff=diff, decay=decay, initial=initial,
                     hold_value=hold, held=box_faces(w, h, d), uptake=uptake,
                     production=production, grid=grid, per_cell_uptake=per_cell_uptake,
                     seed=seed)


def oxygen_disc_200(seed: int = 1000, n: int = 200):
    """200^2 plane, one oxygen field, ~450 Voronoi cells in a radius-60 disc,
    per-cell uptake U[0.5, 1.5]e-3 /s, ring held at 1. The solve uses omega 1.9.
    Another n scales the disc with the edge and the cell count with the area."""
    rng = np.random.default_rng(seed)
    scale = n / 200.0
    ncells = int(round(450 * scale * scale))
    grid = voronoi_disc(n, n, n / 2.0, n / 2.0, 60.0 * scale, ncells, rng)
    rates = np.zeros(ncells + 1)
    rates[1:] = rng.uniform(0.5e-3, 1.5e-3, ncells)
    up = gather(grid, rates)
    return [_field("oxygen", n, n, 1, d_px(2000.0), 0.0, 1.0, 1.0, up, np.zeros(n * n),
                   grid, rates, seed)]


def _plane_fields(n, radius, nseeds, rng, chemokine=True, seed=0):
    c = (n - 1) / 2.0 if n % 2 else n / 2.0
    grid = voronoi_disc(n, n, c, c, radius, nseeds, rng)
    up_rates = np.zeros(nseeds + 1)
    up_rates[1:] = rng.uniform(0.5e-3, 1.5e-3, nseeds)
    flag = rng.random(nseeds) < 0.3
    prod_rates = np.zeros(nseeds + 1)
    prod_rates[1:] = np.where(flag, rng.uniform(0.0, 1e-4, nseeds), 0.0)
    nn = n * n
    fields = [
        _field("oxygen", n, n, 1, d_px(2000.0), 0.0, 1.0, 1.0, gather(grid, up_rates),
               np.zeros(nn), grid, up_rates, seed),
        _field("signal", n, n, 1, d_px(100.0), 1e-4, 0.0, 0.0, np.zeros(nn),
               gather(grid, prod_rates), grid, None, seed),
    ]
    if chemokine:
        tissue = np.flatnonzero(grid != 0)
        pts = rng.choice(tissue, size=64, replace=False)
        prod = np.zeros(nn)
        prod[pts] = CHEMOKINE_STRENGTH
        fields.append(_field("chemokine", n, n, 1, d_px(50.0), 0.0, 0.0, 0.0, np.zeros(nn),
                             prod, grid, None, seed))
    return fields


# Strength of each chemokine point source, c-units per s (the workload leaves it open).
CHEMOKINE_STRENGTH = 1e-3


def multifield_plane(seed: int = 2000, n: int = 1024):
    """1024^2 plane, ~40k Voronoi cells in a radius-450 disc, three fields:
    oxygen (D 2000 um^2/s, uptake U[0.5,1.5]e-3), secreted signal (D 100,
    release U[0,1e-4] in 30% of cells, decay 1e-4), chemokine (D 50, 64 point
    sources). Holds 1, 0, 0 on the ring. Smaller n scales the disc and the cell
    count with the area."""
    rng = np.random.default_rng(seed)
    scale = n / 1024.0
    return _plane_fields(n, 450.0 * scale, max(8, int(round(40000 * scale * scale))), rng,
                         seed=seed)


def spheroid(seed: int = 3000, n: int = 256, ncells: int = 60000, radius: float = 110.0,
            necrotic: float = 40.0):
    """256^3 volume, ~60k Voronoi cells in a radius-110 ball, lognormal per-cell
    uptake (mean 1e-3 /s, sigma 0.3), none within radius 40 of the centre,
    box faces held at 1."""
    rng = np.random.default_rng(seed)
    c = n / 2.0
    grid = voronoi_ball(n, c, radius, ncells, rng)
    sigma = 0.3
    rates = np.zeros(ncells + 1)
    rates[1:] = rng.lognormal(math.log(1e-3) - 0.5 * sigma * sigma, sigma, ncells)
    up = gather(grid, rates)
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    core = ((xx - c) ** 2 + (yy - c) ** 2 + (zz - c) ** 2 < necrotic * necrotic).reshape(-1)
    up[core] = 0.0
    return [_field("oxygen", n, n, n, d_px(2000.0), 0.0, 1.0, 1.0, up, np.zeros(n ** 3),
                   grid, rates, seed)]


def campaign_histories(seed: int = 4000, count: int = 512, lo: int = 128, hi: int = 768):
    """`count` independent plane histories, edge n = round(exp(U[ln lo, ln hi])),
    a centred disc covering U[0.4, 0.8] of the grid, Voronoi cells at the
    density of the 1024^2 multi-field plane, oxygen and signal fields each."""
    rng = np.random.default_rng(seed)
    density = 40000.0 / (math.pi * 450.0 * 450.0)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        f = rng.uniform(0.4, 0.8)
        radius = n * math.sqrt(f / math.pi)
        ncell = max(4, int(round(density * math.pi * radius * radius)))
        sub = np.random.default_rng(seed * 1000 + k)
        out.append(_plane_fields(n, radius, ncell, sub, chemokine=False, seed=seed * 1000 + k))
    return out


def adjoint_spheroid(seed: int = 5000, n: int = 192, ncells: int = 25000, radius: float = 82.0,
            necrotic: float = 40.0):
    """192^3 oxygen volume set up as the 256^3 spheroid, with ~25k cells. Returns the field
    and the target's parameters: D scaled by 1.1 and each cell's uptake by
    (1 + 0.1 U[-1, 1])."""
    (field,) = spheroid(seed=seed, n=n, ncells=ncells, radius=radius, necrotic=necrotic)
    rng = np.random.default_rng(seed + 1)
    noise = 1.0 + 0.1 * rng.uniform(-1.0, 1.0, field.per_cell_uptake.size)
    noise[0] = 1.0
    return field, {"diff_scale": 1.1, "uptake_scale": noise}


def necrotic_mask(n: int, necrotic: float) -> np.ndarray:
    c = n / 2.0
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return ((xx - c) ** 2 + (yy - c) ** 2 + (zz - c) ** 2 < necrotic * necrotic).reshape(-1)


def tissue_box(seed: int = 6000, n: int = 120, radius: float = 45.0):
    """A ball of Voronoi cells in a bath: an n^3 box, cells at the spheroid's
    density (~93 voxels a cell) inside a centred ball, medium (id 0) around it.
    Returns (ids [n^3] u32, number of cells)."""
    rng = np.random.default_rng(seed)
    ncells = max(4, int(round(4.0 / 3.0 * math.pi * radius ** 3 / 93.2)))
    return voronoi_ball(n, n / 2.0, radius, ncells, rng), ncells


def lattice_copies(ids: np.ndarray, w: int, h: int, d: int, count: int, rng) -> np.ndarray:
    """One lattice step in the manner of a Cellular Potts sweep: `count`
    accepted copies, each a voxel taking the id of one of its face neighbours
    where the two differ (cell-cell and cell-medium boundaries move), applied
    in order. Returns the new ids; the input is not changed."""
    out = ids.copy()
    n = w * h * d
    strides = [(1, w), (w, h), (w * h, d)]
    done = 0
    while done < count:
        m = 4 * (count - done) + 64
        v = rng.integers(0, n, m)
        axis = rng.integers(0, 3, m)
        sign = rng.integers(0, 2, m) * 2 - 1
        step = np.array([s for s, _ in strides])[axis]
        extent = np.array([e for _, e in strides])[axis]
        coord = (v // step) % extent
        ok = (coord + sign >= 0) & (coord + sign < extent)
        v, nb = v[ok], (v + sign * step)[ok]
        for a, b in zip(v.tolist(), nb.tolist()):
            if out[a] != out[b]:
                out[a] = out[b]
                done += 1
                if done == count:
                    break
    return out
