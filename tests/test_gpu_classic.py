"""The Classic discretisation (disctype 0) of the FP64 arm against the oracle,
through the C ABI: ClassicStrain's full rate of deformation, model 9's
deviatoric stress, artificial viscosity q (kept between steps and moved with
its particle by the per-step sort) and stress update, ClassicAcceleration.
Bit-identical to the oracle phase by phase and over fused steps (EOS 29; the
Tait EOS 28 evaluates pow, held to 1e-10).  Scenes: the periodic channel
(images carry sigma = q = 0), a small 3D dam break, the lid-driven cavity, a
CSF droplet (normals computed, no CSF force in ClassicAcceleration), a Tait
EOS 28 box.
"""
import dataclasses

import numpy as np
import pytest

from sph07_amd import scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def classic(deck, av_q=0.5):
    deck.disctype = 0
    deck.materials = [dataclasses.replace(m, av_q=av_q) for m in deck.materials]
    return deck


def tait_box():
    d = scenes.cavity(n=24)
    m = d.materials[0]
    c0 = m.eosinput[0]
    d.materials = [dataclasses.replace(m, eos=28, eosinput=(1000.0 * c0 * c0 / 7.0, 7.0, 0.0, 0.0))]
    return d


def squeezed_droplet():
    """The CSF droplet with the fluid moving towards the box centre, so the
    stress and q are non-zero (ClassicAcceleration applies no CSF force)."""
    d = scenes.droplet(n_box=32)
    c = 0.5 * d.x[d.dispbc == 0, :2].max(0)
    v = np.zeros_like(d.x)
    fl = d.dispbc == 0
    v[fl, :2] = -2.0e-2 * (d.x[fl, :2] - c)
    d.v = v
    return d


SCENES = {
    "poiseuille": lambda: classic(scenes.poiseuille(nx=30, ny=30)),
    "dam_break": lambda: classic(scenes.dam_break(nx_fluid=10, ny=8, nz_fluid=10, nx_tank=20, nz_tank=14,
                                                  maxnbr=100)),
    "cavity": lambda: classic(scenes.cavity(n=30)),
    "droplet": lambda: classic(squeezed_droplet()),
    "tait": lambda: classic(tait_box()),
}


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def fields():
    from sph07_amd import native as nv
    return [("x", nv.F_X, orc.F_X), ("v", nv.F_V, orc.F_V), ("a", nv.F_A, orc.F_A), ("rho", nv.F_RHO, orc.F_RHO),
            ("p", nv.F_P, orc.F_P), ("q", nv.F_Q, orc.F_Q), ("sigma_diag", nv.F_SIGMA_DIAG, orc.F_SIGMA_DIAG),
            ("sigma_off", nv.F_SIGMA_OFF, orc.F_SIGMA_OFF)]


def same(g, o, exact, what):
    for name, fg, fo in fields():
        a, b = g.field(fg), o.field(fo)
        if exact:
            assert np.array_equal(a, b), f"{what}: {name} differs (rel {rel(a, b):.3e})"
        else:
            assert rel(a, b) <= RTOL, f"{what}: {name} rel {rel(a, b):.3e}"


@pytest.mark.parametrize("name", list(SCENES))
def test_classic_phases(name):
    from sph07_amd import native as nv
    deck = SCENES[name]()
    print(f"classic phases: {name}, {deck.np} particles", flush=True)
    exact = name != "tait"
    o, g = orc.Oracle(deck), nv.Scene(deck)
    checks = {
        orc.PH_STRAIN: [("trd", nv.F_TRD, orc.F_TRD), ("rod_diag", nv.F_ROD_DIAG, orc.F_ROD_DIAG),
                        ("rod_off", nv.F_ROD_OFF, orc.F_ROD_OFF), ("x", nv.F_X, orc.F_X)],
        orc.PH_NORMALS: [("kappa", nv.F_KAPPA, orc.F_KAPPA)] if deck.surface_tension else [],
        orc.PH_CONSTITUTIVE: [("rho", nv.F_RHO, orc.F_RHO), ("p", nv.F_P, orc.F_P), ("q", nv.F_Q, orc.F_Q),
                              ("sigma_diag", nv.F_SIGMA_DIAG, orc.F_SIGMA_DIAG),
                              ("sigma_off", nv.F_SIGMA_OFF, orc.F_SIGMA_OFF)],
        orc.PH_MOMENTUM: [("a", nv.F_A, orc.F_A)],
        orc.PH_VELOCITY: [("v", nv.F_V, orc.F_V)],
        orc.PH_MOVE: [("x", nv.F_X, orc.F_X)],
    }
    for step in range(3):
        for ph in range(orc.NPHASES):
            o.phase(ph)
            g.phase(ph)
            for fname, fg, fo in checks.get(ph, []):
                a, b = g.field(fg), o.field(fo)
                if exact:
                    assert np.array_equal(a, b), f"step {step} phase {ph} {fname}: rel {rel(a, b):.3e}"
                else:
                    assert rel(a, b) <= RTOL, f"step {step} phase {ph} {fname}: rel {rel(a, b):.3e}"
        go, gg = o.globals(), g.globals()
        for k in ("dt", "dtold", "ptime", "critts"):
            assert gg[k] == go[k] or abs(gg[k] - go[k]) <= RTOL * abs(go[k]), (step, k, gg[k], go[k])


@pytest.mark.parametrize("name,steps", [("poiseuille", 100), ("dam_break", 40), ("cavity", 100),
                                        ("droplet", 60), ("tait", 60)])
def test_classic_fused_steps(name, steps):
    from sph07_amd import native as nv
    deck = SCENES[name]()
    print(f"classic fused: {name}, {deck.np} particles, {steps} steps", flush=True)
    o, g = orc.Oracle(deck), nv.Scene(deck)
    assert o.run(steps) == g.Solution(steps) == steps
    # the artificial viscosity was switched on somewhere (compression)
    assert np.abs(o.field(orc.F_Q)).max() > 0
    same(g, o, exact=name != "tait", what=f"{name} after {steps} steps")
    go, gg = o.globals(), g.globals()
    for k in ("dt", "ptime"):
        assert abs(gg[k] - go[k]) <= RTOL * abs(go[k]), (k, gg[k], go[k])


def test_classic_launch_shapes_same_bits():
    from sph07_amd import native as nv
    deck = SCENES["dam_break"]()
    ref, other = nv.Scene(deck), nv.Scene(deck)
    ref.set_launch_shape()
    other.set_launch_shape(strain_batch=8, momentum_batch=8, cell_order="row-major")
    assert ref.Solution(20) == other.Solution(20) == 20
    for name, fg, _ in fields():
        assert np.array_equal(ref.field(fg), other.field(fg)), name


def test_classic_unsupported_combinations():
    from sph07_amd import native as nv
    with pytest.raises(nv.SphError) as e:
        nv.Scene(SCENES["cavity"](), nv.ARM_FP32C)
    assert e.value.code == 4
    cyl = scenes.cylinder(nx=30, ny=40, dx=5e-4, radius=3e-3, centre=(0.008, 0.01))
    cyl.disctype = 0
    with pytest.raises(nv.SphError) as e:
        nv.Scene(cyl)
    assert e.value.code == 4
    b = nv.Batch(2, nv.ARM_FP64)
    with pytest.raises(nv.SphError) as e:
        b.add(SCENES["cavity"]())
    assert e.value.code == 4
    b.close()
