"""The library's EnSight CASE output (sph_output.cpp) byte for byte against the
reference's writer restated in Python (tests/ensight_ref.py, from
MOutput.cpp:116-710), fed the fields downloaded at the same moment: the
file set, the block structure, the line breaks (pairs of nodes, six values,
WriteCase3D's breaks on the particle number), the case file's time list and
every printed digit.  1D, 2D with two materials, 3D with two materials and
the Classic discretisation's full stress.

The quantities the device path does not evolve (efps, energy, temperature,
the boundary flag) are fed as zero, and the Morris discretisation's stress as
-p I, the values the library documents for them."""
import dataclasses
import os

import numpy as np
import pytest

from ensight_ref import EnsightModel, VAR_FILES

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def line_1d(n=40, layers=3, dx=1.0e-4):
    from sph07_amd.deck import Deck, Material
    i = np.arange(-layers, n + layers)
    x = np.zeros((i.size, 3))
    x[:, 0] = (i + 0.5) * dx
    wall = (i < 0) | (i >= n)
    v = np.zeros_like(x)
    v[~wall, 0] = 1.0e-3 * np.sin(2.0 * np.pi * x[~wall, 0] / (n * dx))
    mat = Material(eos=29, visc_type=2, rho=1000.0, av_l=1.0e-6, mass=1000.0 * dx * i.size, h=1.3 * dx,
                   eosinput=(0.1, 0.0, 0.0, 0.0), av_alpha=0.1)
    return Deck(ndim=1, x=x, v=v, mat_id=np.ones(i.size, np.int32), materials=[mat],
                dispbc=np.where(wall, 7, 0).astype(np.int32), tcrit_opt=3, maxnbr=20, name="line_1d")


def two_material_dam():
    """A small 3D dam break whose walls are a second material (the 3D
    writer's line breaks on the particle number show only with a particle
    count per material that is not a multiple of six)."""
    from sph07_amd import scenes
    d = scenes.dam_break(nx_fluid=7, ny=5, nz_fluid=7, nx_tank=13, nz_tank=10, maxnbr=100)
    wall = d.dispbc != 0
    d.materials = [d.materials[0], dataclasses.replace(d.materials[0])]
    d.mat_id = np.where(wall, 2, 1).astype(np.int32)
    return d


def classic_cavity():
    from sph07_amd import scenes
    d = scenes.cavity(n=20)
    d.disctype = 0
    d.materials = [dataclasses.replace(m, av_q=0.5) for m in d.materials]
    return d


def kuglica():
    from sph07_amd.native import read_mcm
    return read_mcm(os.path.join(GOLD, "kuglica.mcm"))


CASES = {"1d": line_1d, "2d_kuglica": kuglica, "3d_two_materials": two_material_dam,
         "2d_classic": classic_cavity}


def downloads(g, n):
    from sph07_amd import native as nv
    sd, so = g.field(nv.F_SIGMA_DIAG), g.field(nv.F_SIGMA_OFF)
    z = np.zeros(n)
    return {"x": g.field(nv.F_X), "v": g.field(nv.F_V), "a": g.field(nv.F_A), "mass": g.field(nv.F_MASS),
            "boundary": z, "nnbr": g.field(nv.F_NNBR), "rho": g.field(nv.F_RHO), "efps": z, "energy": z,
            "c": g.field(nv.F_C), "temper": z, "sigma": np.stack([sd[:, 0], sd[:, 1], so[:, 0]], axis=1)}


@pytest.mark.parametrize("name", list(CASES))
def test_ensight_bytes(tmp_path, name):
    from sph07_amd import native as nv
    from sph07_amd.native import Scene, StateOutput
    deck = CASES[name]()
    n = deck.np
    print(f"output format: {name}, {n} particles, ndim {deck.ndim}", flush=True)
    g = Scene(deck)
    base = str(tmp_path / name)
    out = StateOutput(g, base, deck)
    model = EnsightModel(base, deck.ndim, deck.mat_id, len(deck.materials), g.field(nv.F_X))
    for steps in (0, 9, 6, 11):
        if steps:
            assert g.Solution(steps) == steps
        out.state()
        model.state(downloads(g, n), g.globals()["ptime"])
    out.close()
    g.close()
    if name == "2d_classic":
        sig_off = [float(ln[36:48]) for ln in model.files["_sig.ens"].split("\n") if len(ln) == 72]
        assert any(v != 0.0 for v in sig_off), "the Classic stress has a shear component"
    for ext in [".geo", ".case", "_time.txt"] + [e for e, _ in VAR_FILES]:
        got = open(base + ext).read()
        want = model.files[ext]
        if got != want:
            a, b = got.split("\n"), want.split("\n")
            k = next((i for i, (p, q) in enumerate(zip(a, b)) if p != q), min(len(a), len(b)))
            raise AssertionError(f"{name}{ext}: line {k + 1} differs: {a[k:k + 1]!r} vs {b[k:k + 1]!r} "
                                 f"({len(a)} vs {len(b)} lines)")
