"""The synthetic inputs: seeded, and shaped as their docstrings say."""
import numpy as np

from cellweave_amd import cases
from cellweave_amd import workloads as wl


def neighbour_ids(ids, w, h, d, i):
    x, y, z = i % w, (i // w) % h, i // (w * h)
    out = set()
    for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        a, b, c = x + dx, y + dy, z + dz
        if 0 <= a < w and 0 <= b < h and 0 <= c < d:
            out.add(int(ids[(c * h + b) * w + a]))
    return out


def test_lattice_copies_move_boundaries_to_a_neighbours_id():
    ids, ncells = cases.tissue_box(seed=17, n=24, radius=9.0)
    assert ncells == round(4.0 / 3.0 * np.pi * 9.0 ** 3 / 93.2)
    assert set(np.unique(ids)) <= set(range(ncells + 1))
    new = cases.lattice_copies(ids, 24, 24, 24, 40, np.random.default_rng(3))
    again = cases.lattice_copies(ids, 24, 24, 24, 40, np.random.default_rng(3))
    assert np.array_equal(new, again)
    changed = np.flatnonzero(new != ids)
    assert 0 < changed.size <= 40
    for i in changed.tolist():
        # each copy takes a face neighbour's id at the time it is made
        assert int(new[i]) in neighbour_ids(ids, 24, 24, 24, i) | neighbour_ids(new, 24, 24, 24, i)


def test_tissue_steps_are_seeded_chains():
    seq, ncells = wl.tissue_steps(n=20, radius=7.0, steps=3, copies=10, seed=9)
    seq2, _ = wl.tissue_steps(n=20, radius=7.0, steps=3, copies=10, seed=9)
    assert len(seq) == 4
    for a, b in zip(seq, seq2):
        assert np.array_equal(a, b)
    for a, b in zip(seq, seq[1:]):
        assert 0 < int((a != b).sum()) <= 10
