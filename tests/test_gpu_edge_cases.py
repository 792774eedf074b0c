"""Edge cases of the step loop against the oracle, through the C ABI:
the smallest scenes (one and two particles), a 1D scene, the end-time stop
of MSolution::Solution and a grid that outgrows its cell arrays (the host
grows them and runs the step again).  Scenes without a rigid body are held
to bit-identity, as in the parity tests.
"""
import math

import numpy as np
import pytest

from sph07_amd.deck import Deck, Material
from sph07_amd import scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = [("x", orc.F_X), ("v", orc.F_V), ("a", orc.F_A), ("rho", orc.F_RHO), ("p", orc.F_P)]


def assert_identical(g, o, what):
    for name, f in FIELDS:
        a, b = g.field(f), o.field(f)
        assert np.array_equal(a, b), f"{what}: {name} differs by {np.abs(a - b).max():.3e}"
    go, gg = o.globals(), g.globals()
    for k in ("dt", "ptime", "timestep"):
        assert gg[k] == go[k], (what, k, gg[k], go[k])


def free_deck(x, v, ndim, dx, h_over_dx=1.3, maxnbr=40, gravity=True):
    """Fluid particles of one material (linear EOS, Morris viscosity) in free
    space, with gravity along the last axis when asked."""
    n = x.shape[0]
    rho0, c0 = 1000.0, 1.0
    mat = Material(eos=29, rho=rho0, av_l=1.0e-6, mass=rho0 * dx ** ndim * n, h=h_over_dx * dx,
                   eosinput=(c0, 0.0, 0.0, 0.0))
    d = Deck(ndim=ndim, x=x, v=v, mat_id=np.ones(n, np.int32), materials=[mat],
             dispbc=np.zeros(n, np.int32), tcrit_opt=3, maxnbr=maxnbr, baseaccel=1 if gravity else 0,
             name="free")
    if gravity:
        d.base_a[ndim - 1] = -9.81
    return d


def run_both(deck, steps):
    from sph07_amd.native import Scene
    o = orc.Oracle(deck)
    g = Scene(deck)
    nn_o, ids_o = o.neighbours()
    nn_g, ids_g = g.neighbours()
    assert np.array_equal(nn_o, nn_g)
    for i in range(deck.np):
        assert np.array_equal(ids_o[i, :nn_o[i]], ids_g[i, :nn_g[i]]), i
    n_o, n_g = o.run(steps), g.Solution(steps)
    assert n_o == n_g, (n_o, n_g)
    return o, g, n_g


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_lone_particle(ndim):
    """One particle under gravity: no neighbours, the critical step from the
    acceleration alone."""
    x = np.zeros((1, 3))
    deck = free_deck(x, np.zeros((1, 3)), ndim, 1.0e-3)
    o, g, n = run_both(deck, 25)
    assert n == 25
    assert_identical(g, o, f"lone particle {ndim}D")


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_particle_pair(ndim):
    """Two particles inside each other's support, approaching."""
    dx = 1.0e-3
    x = np.zeros((2, 3))
    x[1, 0] = dx
    v = np.zeros((2, 3))
    v[0, 0], v[1, 0] = 1.0e-3, -1.0e-3
    deck = free_deck(x, v, ndim, dx, gravity=False)
    o, g, n = run_both(deck, 40)
    assert n == 40
    nn, _ = g.neighbours()
    assert nn.tolist() == [1, 1]
    assert_identical(g, o, f"pair {ndim}D")


def line_1d(n=200, layers=3, dx=1.0e-4, amp=1.0e-3):
    """1D column of fluid between fixed wall particles, started with a
    sinusoidal velocity."""
    i = np.arange(-layers, n + layers)
    x = np.zeros((i.size, 3))
    x[:, 0] = (i + 0.5) * dx
    wall = (i < 0) | (i >= n)
    v = np.zeros_like(x)
    L = n * dx
    v[~wall, 0] = amp * np.sin(2.0 * math.pi * x[~wall, 0] / L)
    rho0, c0 = 1000.0, 0.1
    mat = Material(eos=29, visc_type=2, rho=rho0, av_l=1.0e-6, mass=rho0 * dx * i.size, h=1.3 * dx,
                   eosinput=(c0, 0.0, 0.0, 0.0), av_alpha=0.1)
    d = Deck(ndim=1, x=x, v=v, mat_id=np.ones(i.size, np.int32), materials=[mat],
             dispbc=np.where(wall, 7, 0).astype(np.int32), tcrit_opt=3, maxnbr=20, name="line_1d")
    return d


def test_one_dimensional_column():
    """The 1D kernels (3-cell stencil) with walls and the Monaghan term: bit
    identical to the oracle after 150 steps; the FP32c arm within 1e-5."""
    from sph07_amd.native import Scene, ARM_FP32C, F_V
    deck = line_1d()
    o, g, n = run_both(deck, 150)
    assert n == 150
    assert_identical(g, o, "1D column")
    g32 = Scene(deck, ARM_FP32C)
    assert g32.Solution(150) == 150
    moving = deck.dispbc == 0
    v64, v32 = g.field(F_V)[moving], g32.field(F_V)[moving]
    err = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    assert err <= 1e-5, err


def test_end_time_stops_the_loop():
    """Solution stops where the reference's stop test (t - dt >= endtime) does:
    the same number of steps and the same state as the oracle."""
    deck = scenes.poiseuille(nx=20, ny=20)
    probe = orc.Oracle(deck)
    probe.run(30)
    deck.endtime = probe.globals()["ptime"]          # reached in about 30 steps
    o, g, n = run_both(deck, 10000)
    assert 20 < n < 10000, n
    assert_identical(g, o, "end time")
    assert g.globals()["done"] == 1


def test_grid_outgrows_the_cell_arrays():
    """A particle leaving a cavity at speed stretches the bounding box, and
    with it the 2h grid, far beyond the cell arrays sized at creation; the
    step that finds the grid too large is repeated on grown arrays.  The run
    stays bit-identical to the oracle."""
    base = scenes.cavity(n=20)
    dx = base.meta["dx"]
    x = np.concatenate([base.x, [[-2.0 * base.meta["L"], 0.5 * base.meta["L"], 0.0]]])
    v = np.concatenate([base.v, [[-2.0, 0.0, 0.0]]])
    mat = base.materials[0]
    mat.mass = mat.rho * dx * dx * x.shape[0]
    deck = Deck(ndim=2, x=x, v=v, mat_id=np.ones(x.shape[0], np.int32), materials=[mat],
                dispbc=np.concatenate([base.dispbc, [0]]).astype(np.int32), tcrit_opt=3,
                maxnbr=base.maxnbr, name="cavity_escape")
    o, g, n = run_both(deck, 40)
    assert n == 40
    assert_identical(g, o, "grid growth")
    # the grid at the end against the cell arrays sized at creation (the
    # code space is a power of two per axis; 1.25 x + 4096 spare)
    h2 = 2.0 * mat.h

    def code_space(xs):
        ext = xs.max(axis=0)[:2] - xs.min(axis=0)[:2]
        cells = np.floor(ext / h2).astype(int) + 2
        return int(np.prod([1 << int(math.ceil(math.log2(c))) for c in cells]))

    start = code_space(x)
    assert code_space(g.field(orc.F_X)) > start * 1.25 + 4096
