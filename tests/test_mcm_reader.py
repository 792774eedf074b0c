"""The .mcm deck reader (sph_deck_read, MSimulationData::GetInput
MSimulationData.cpp:186-1068) on two of the reference's own decks, on
decks written by sph07_amd.mcm, and on malformed decks.  CPU only."""
import os

import numpy as np
import pytest

from sph07_amd import scenes
from sph07_amd.mcm import write_mcm
from sph07_amd.native import SphError, read_mcm

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF = "/root/reference"


def test_rigidbody_deck():
    d = read_mcm(os.path.join(GOLD, "rigidbody.mcm"))
    assert (d.np, d.ndim, len(d.materials), d.disctype, d.tcrit_opt, d.maxnbr) == (1891, 2, 2, 2, 3, 40)
    assert d.rigid_body == 1 and d.rb_id == 2
    assert d.rb_total_mass == 2.65e-4 and d.rb_inertia == 2.25e-12
    # material cards after CheckDefaults; av_l is the viscosity column
    assert d.materials[0].av_l == 1.0e-6 and d.materials[0].eos == 29
    # base acceleration masked by nthp (only y is switched on)
    assert d.baseaccel == 1 and d.base_a[0] == 0.0 and d.base_a[1] == 0.02
    # boundary card present but all codes zero -> boundary off (:966-967)
    assert d.boundary == 0
    assert set(np.unique(d.mat_id)) == {1, 2}
    assert d.x[0, 0] == -5.35714e-04


def test_kuglica_deck_two_colours():
    d = read_mcm(os.path.join(GOLD, "kuglica.mcm"))
    assert d.surface_tension == 1 and d.rigid_body == 0
    assert len(np.unique(d.color)) == 2
    assert d.materials[1].surface_tension == 2e-07


# the reference's decks kept as parity fixtures: what their cards state
# (ndim, particles, materials, surface tension, rigid body, periodic boundary card, EOS)
FIXTURE_DECKS = {
    "stenosis.mcm": (2, 1891, 1, 0, 0, 1, [28]),
    "rbcell.mcm": (2, 1891, 2, 1, 0, 1, [29, 29]),
    "cylinders.mcm.gz": (2, 9044, 1, 0, 0, 0, [28]),
    "flowpast.mcm.gz": (2, 10201, 1, 0, 0, 1, [29]),
    "decuzzi.mcm.gz": (2, 18411, 2, 0, 1, 1, [29, 29]),
    "rbcell3d.mcm.gz": (3, 43005, 2, 1, 0, 1, [29, 29]),
}


@pytest.mark.parametrize("name", list(FIXTURE_DECKS))
def test_fixture_decks_parse(name, tmp_path):
    import gzip
    path = os.path.join(GOLD, name)
    if name.endswith(".gz"):
        out = tmp_path / name[:-3]
        with gzip.open(path, "rb") as f:
            out.write_bytes(f.read())
        path = str(out)
    d = read_mcm(path)
    nd, n, nmat, st, rb, bnd, eos = FIXTURE_DECKS[name]
    assert (d.ndim, d.np, len(d.materials), d.surface_tension, d.rigid_body, d.boundary) == (nd, n, nmat, st, rb,
                                                                                             bnd)
    assert [m.eos for m in d.materials] == eos
    assert d.disctype == 2 and d.tcrit_opt == 3
    assert set(np.unique(d.mat_id)) == set(range(1, nmat + 1))


@pytest.mark.parametrize("gen", ["poiseuille", "cavity", "droplet", "cylinder"])
def test_write_read_round_trip(tmp_path, gen):
    kw = {"poiseuille": dict(nx=12, ny=12), "cavity": dict(n=10), "droplet": dict(n_box=16),
          "cylinder": dict(nx=20, ny=30, dx=1e-3, radius=4e-3, centre=(0.01, 0.015))}[gen]
    d = scenes.by_name(gen, **kw)
    p = str(tmp_path / f"{gen}.mcm")
    write_mcm(d, p)
    r = read_mcm(p)
    assert r.np == d.np and r.ndim == d.ndim
    assert np.allclose(r.x, d.x, rtol=1e-12, atol=0)
    assert np.array_equal(r.dispbc, d.dispbc) and np.array_equal(r.mat_id, d.mat_id)
    assert np.array_equal(r.color, d.color)
    assert np.allclose(r.v, d.v, rtol=1e-3, atol=0)
    assert np.array_equal(r.boundary_code, d.boundary_code)
    assert r.rigid_body == d.rigid_body and r.surface_tension == d.surface_tension
    for a, b in zip(r.materials, d.materials):
        assert abs(a.h - b.h) <= 1e-3 * b.h and abs(a.av_l - b.av_l) <= 1e-3 * b.av_l


def test_out_of_order_node_is_an_error(tmp_path):
    d = scenes.poiseuille(nx=4, ny=4)
    p = tmp_path / "bad.mcm"
    write_mcm(d, str(p))
    lines = p.read_text().splitlines()
    k = next(i for i, s in enumerate(lines) if s.startswith("       3"))
    lines[k] = "       9" + lines[k][8:]
    p.write_text("\n".join(lines) + "\n")
    with pytest.raises(SphError, match="out of order"):
        read_mcm(str(p))


def test_missing_file():
    with pytest.raises(SphError, match="Cannot open"):
        read_mcm("/nonexistent/deck.mcm")


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference decks not present")
def test_reference_decks_parse():
    ok = []
    for f in sorted(os.listdir(REF)):
        if not f.endswith(".mcm") or f in ("1d_sod.mcm", "2d_block_e.mcm"):
            continue   # these two use an older card layout
        try:
            d = read_mcm(os.path.join(REF, f))
            assert d.np > 0
            ok.append(f)
        except SphError as e:
            assert "Lennard-Jones" in str(e), (f, str(e))
    assert len(ok) >= 15
