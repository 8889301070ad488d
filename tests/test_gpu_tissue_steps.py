"""The reference's per-lattice-step loop end to end, GPU against the oracle.

simulation.rs:1068-1083 settles the oxygen field once to 1e-8; every lattice
step then declares the sources (uptake of the living cells, release, the
medium bath held at its level; simulation.rs:2138-2145), relaxes from the last
field to 0.01 x uptake within 20,000 sweeps (simulation.rs:1144-1154) and
reads the per-cell means (simulation.rs:1157). With every array on the device,
each step must give the oracle's sweep count, values, held mask and means bit
for bit."""
import numpy as np
import pytest

from cellweave_amd import cases
from cellweave_amd import workloads as wl
from cellweave_amd.field import Field, Lattice
from helpers import bits_equal
from oracle.orf import OracleField

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("n,radius,copies", [(40, 15.0, 150), (33, 11.0, 60)])
def test_lattice_steps_match_the_oracle_bit_for_bit(n, radius, copies):
    say(f"case tissue steps n={n} radius={radius} copies={copies}")
    seq, ncells = wl.tissue_steps(n=n, radius=radius, steps=5, copies=copies, seed=6100 + n)
    consumes = np.ones(ncells + 1, np.uint8)
    consumes[0] = 0
    consumes[3] = 0  # one cell that does not consume
    diff = cases.d_px(2000.0)
    lat = Lattice(n, n, n)
    g = Field(n, n, n, diff, 0.0, 1.0)
    ref = OracleField(n, n, n, diff, 0.0, 1.0)

    def declare(ids):
        lat.upload(ids)
        g.apply_uptake(lat, consumes, wl.TISSUE_UPTAKE)
        g.release_all()
        g.hold_medium_bath(lat, 1.0)
        ref.apply_uptake(ids, consumes, wl.TISSUE_UPTAKE)
        ref.release_all()
        ref.hold_bath_at(1.0, (ids == 0).astype(np.uint8))

    declare(seq[0])
    assert g.relax(*wl.TISSUE_WARMUP) == ref.relax(*wl.TISSUE_WARMUP)
    assert bits_equal(g.download(), ref.values)
    for k, ids in enumerate(seq[1:]):
        declare(ids)
        want = ref.relax(wl.TISSUE_TOLERANCE * wl.TISSUE_UPTAKE, wl.TISSUE_MAX_SWEEPS)
        got = g.relax(wl.TISSUE_TOLERANCE * wl.TISSUE_UPTAKE, wl.TISSUE_MAX_SWEEPS)
        say(f"  step {k}: sweeps gpu {got} oracle {want}")
        assert got == want
        assert bits_equal(g.download(), ref.values)
        held, _, _ = g.state()
        assert np.array_equal(held.astype(bool), ref.held.astype(bool))
        assert bits_equal(g.mean_per_cell(lat, ncells + 1), ref.mean_per_cell(ids, ncells + 1))
