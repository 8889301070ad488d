"""GPU adjoint gradients against the CPU FP64 adjoint and central finite
differences (forward solves in FP64 on the GPU, converged to 1e-14)."""
import numpy as np
import pytest

from cellweave_amd.field import CRIT_REL_RESIDUAL, Field
from oracle.orf import OracleField

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def lattice_case(w, h, d, ncells, seed):
    rng = np.random.default_rng(seed)
    n = w * h * d
    grid = rng.integers(1, ncells, size=n).astype(np.uint32)
    faces = np.zeros((d, h, w), np.uint8)
    faces[:, 0, :] = faces[:, -1, :] = faces[:, :, 0] = faces[:, :, -1] = 1
    if d > 1:
        faces[0] = faces[-1] = 1
    faces = faces.reshape(-1)
    rates = np.concatenate([[0.0], rng.uniform(0.5e-3, 1.5e-3, ncells - 1)])
    target = rng.uniform(0.85, 1.0, n)
    return grid, faces, rates, target


def gpu_solve(w, h, d, diff, g, rates, grid, faces, omega):
    f = Field(w, h, d, diff, 0.0, 1.0)
    up = np.where(faces == 1, 0.0, rates[grid])
    vals = np.where(faces == 1, g, 1.0)
    f.upload(values=vals, held=faces, uptake=up)
    f.relax_ex(omega=omega, criterion=CRIT_REL_RESIDUAL, tolerance=1e-14, max_sweeps=500_000,
               sweeps_per_pass=1)
    return f


@pytest.mark.parametrize("shape", [(48, 40, 1), (20, 18, 16)])
def test_adjoint_matches_the_cpu_adjoint(shape):
    w, h, d = shape
    say(f"case adjoint vs cpu adjoint {shape}")
    ncells = 12
    grid, faces, rates, target = lattice_case(w, h, d, ncells, sum(shape))
    omega = 1.7
    f = gpu_solve(w, h, d, 0.3, 1.0, rates, grid, faces, omega)
    c = f.download()
    got = f.adjoint(target, grid=grid, ncells=ncells, tolerance=1e-13, omega=omega)
    ref = OracleField(w, h, d, 0.3, 0.0, 1.0)
    ref.values[:] = c
    ref.uptake[:] = np.where(faces == 1, 0.0, rates[grid])
    ref.held[:] = faces
    want = ref.adjoint(target, omega=omega, rel_tol=1e-13, grid=grid, ncells=ncells)
    assert got["rel_residual"] <= 1e-13
    assert abs(got["dl_dd"] - want["dD"]) <= 1e-9 * abs(want["dD"])
    assert abs(got["dl_dg"] - want["dg"]) <= 1e-9 * abs(want["dg"])
    scale = np.max(np.abs(want["du"]))
    assert np.max(np.abs(got["du"] - want["du"])) <= 1e-9 * scale
    assert abs(got["loss"] - 0.5 * np.sum((c - target) ** 2)) <= 1e-12 * got["loss"]


def test_adjoint_matches_finite_differences():
    w = h = d = 18
    say("case adjoint vs central differences 18^3")
    ncells = 10
    grid, faces, rates, target = lattice_case(w, h, d, ncells, 99)
    omega, diff = 1.75, 0.3

    def loss(diff_, g_, rates_):
        f = gpu_solve(w, h, d, diff_, g_, rates_, grid, faces, omega)
        return 0.5 * np.sum((f.download() - target) ** 2)

    f = gpu_solve(w, h, d, diff, 1.0, rates, grid, faces, omega)
    adj = f.adjoint(target, grid=grid, ncells=ncells, tolerance=1e-14, omega=omega)
    hD = 1e-6 * diff
    fd = (loss(diff + hD, 1.0, rates) - loss(diff - hD, 1.0, rates)) / (2 * hD)
    assert abs(adj["dl_dd"] - fd) <= 1e-6 * abs(fd), (adj["dl_dd"], fd)
    fd = (loss(diff, 1.0 + 1e-6, rates) - loss(diff, 1.0 - 1e-6, rates)) / 2e-6
    assert abs(adj["dl_dg"] - fd) <= 1e-6 * abs(fd), (adj["dl_dg"], fd)
    for k in (1, 4, 9):
        hk = 1e-6 * rates[k]
        up, dn = rates.copy(), rates.copy()
        up[k] += hk
        dn[k] -= hk
        fd = (loss(diff, 1.0, up) - loss(diff, 1.0, dn)) / (2 * hk)
        assert abs(adj["du"][k] - fd) <= 1e-5 * abs(fd), (k, adj["du"][k], fd)


def test_adjoint_refuses_a_floored_field():
    say("case adjoint floored")
    n = 40
    yy, xx = np.mgrid[0:n, 0:n]
    disc = ((xx - 20) ** 2 + (yy - 20) ** 2 <= 15 ** 2).reshape(-1)
    f = Field(n, n, 1, 0.2, 0.0, 1.0)
    f.upload(values=np.ones(n * n), held=(~disc).astype(np.uint8),
             uptake=np.where(disc, 0.02, 0.0))
    f.relax_ex(omega=1.9, criterion=CRIT_REL_RESIDUAL, tolerance=1e-10, max_sweeps=200_000)
    with pytest.raises(ValueError, match="floor is active"):
        f.adjoint(np.ones(n * n))


def test_spheroid_adjoint_matches_cpu_and_differences():
    """The 192^3 adjoint workload scaled to 48^3: Voronoi spheroid with a
    necrotic core, forward in FP32+refinement, adjoint in FP64, checked against
    the CPU FP64 adjoint (1e-9) and central differences at h = 1e-6 relative
    (1e-6), the perturbed solves in correction form."""
    from cellweave_amd.field import PREC_FP32_REFINE
    from cellweave_amd.workloads import SpheroidAdjoint

    n = 48
    s = n / 192
    say(f"case spheroid adjoint {n}^3")
    case = SpheroidAdjoint(n=n, ncells=int(25000 * s ** 3), radius=82 * s, necrotic=40 * s)
    case.prepare()
    rates = case.rates
    f, _ = case.solve(case.fc.diff, 1.0, rates, precision=PREC_FP32_REFINE, tol=1e-12)
    c = f.download()
    assert c.min() > 0
    assert np.max(np.abs(c - case.base)) <= 1e-10
    got = f.adjoint(case.target, grid=case.fc.grid, ncells=rates.size, tolerance=1e-12,
                    omega=case.omega)
    ref = OracleField(n, n, n, case.fc.diff, 0.0, 1.0)
    ref.values[:] = c
    ref.uptake[:] = case.uptake_of(rates)
    ref.held[:] = case.fc.held
    want = ref.adjoint(case.target, omega=case.omega, rel_tol=1e-12, grid=case.fc.grid,
                       ncells=rates.size)
    assert abs(got["dl_dd"] - want["dD"]) <= 1e-9 * abs(want["dD"])
    assert abs(got["dl_dg"] - want["dg"]) <= 1e-9 * abs(want["dg"])
    assert np.max(np.abs(got["du"] - want["du"])) <= 1e-9 * np.max(np.abs(want["du"]))
    live = np.flatnonzero(np.abs(got["du"]) > 0)
    cells = np.random.default_rng(48).choice(live, size=4, replace=False)
    fd = case.differences(h=1e-6, cells=cells)
    assert abs(got["dl_dd"] - fd["D"]) <= 1e-6 * abs(fd["D"]), (got["dl_dd"], fd["D"])
    assert abs(got["dl_dg"] - fd["g"]) <= 1e-6 * abs(fd["g"]), (got["dl_dg"], fd["g"])
    for k, v in fd["cells"].items():
        assert abs(got["du"][k] - v) <= 1e-6 * abs(v), (k, got["du"][k], v)
