"""The captured step sequence of sph_solution (HIP graphs of 16 steps) and the
launch-shape tuner leave results unchanged.

A step moves the particle state from one buffer set to the other, so a graph
captured from one set is only valid when launched from that set.  Runs that
end a call on an odd step (a direct-launch remainder) or that mix the phase
entry points with sph_solution must give the bits of one uninterrupted run.
"""
import numpy as np
import pytest

from sph07_amd import scenes

pytestmark = pytest.mark.gpu


def _fields():
    from sph07_amd.native import F_X, F_V, F_A, F_RHO, F_P
    return (("x", F_X), ("v", F_V), ("a", F_A), ("rho", F_RHO), ("p", F_P))


def _same(s1, s2, what):
    for name, f in _fields():
        a, b = s1.field(f), s2.field(f)
        assert np.array_equal(a, b), f"{what}: {name} differs (max {np.abs(a - b).max():.3e})"
    g1, g2 = s1.globals(), s2.globals()
    for k in ("dt", "dtold", "ptime", "critts", "timestep"):
        assert g1[k] == g2[k], (what, k, g1[k], g2[k])


@pytest.mark.parametrize("arm_name", ["fp64", "fp32c"])
def test_odd_remainder_then_graph(arm_name):
    """Solution(17) (one graph and one direct step, ending on the other buffer
    set) followed by Solution(16) (a graph launched from that set) equals
    Solution(33)."""
    from sph07_amd.native import Scene, ARM_FP64, ARM_FP32C
    arm = ARM_FP64 if arm_name == "fp64" else ARM_FP32C
    deck = scenes.poiseuille(nx=40, ny=40)
    split, whole = Scene(deck, arm), Scene(deck, arm)
    print(f"graph parity: poiseuille {deck.np} particles, arm {arm_name}", flush=True)
    assert split.Solution(17) == 17
    assert split.Solution(16) == 16
    assert split.Solution(1) == 1
    assert split.Solution(32) == 32
    assert whole.Solution(66) == 66
    _same(split, whole, f"{arm_name} 17+16+1+32 vs 66")


def test_phases_then_graph():
    """A step through the phase entry points (MSolution's phase surface),
    then graph-launched steps, twice over, equals one uninterrupted run."""
    from sph07_amd.native import Scene
    deck = scenes.dam_break(nx_fluid=10, ny=8, nz_fluid=10, nx_tank=20, nz_tank=14, maxnbr=100)
    print(f"graph parity: dam break {deck.np} particles, phases then graph", flush=True)
    mixed, whole = Scene(deck), Scene(deck)
    for ph in range(9):
        mixed.phase(ph)
    assert mixed.Solution(16) == 16
    for ph in range(9):
        mixed.phase(ph)
    assert mixed.Solution(32) == 32
    assert whole.Solution(50) == 50
    _same(mixed, whole, "phase step + 16 + phase step + 32 vs 50")


def _acceleration_bound_channel():
    """A periodic channel (10,600 particles) whose critical step is set by the
    acceleration term of tcrit option 3, 0.25 sqrt(h/|a|)
    (MSolution.cpp:69-70): a body force of 1 m/s^2 with a sound speed of
    1e-3 m/s and nu = 1e-9 m^2/s puts the sound and viscous limits above it."""
    import dataclasses
    deck = scenes.poiseuille(nx=100, ny=100, dx=1.0e-5, nu=1.0e-9, force=1.0)
    m = deck.materials[0]
    deck.materials = [dataclasses.replace(m, eosinput=(1.0e-3, 0.0, 0.0, 0.0))]
    return deck


@pytest.mark.parametrize("arm_name", ["fp64", "fp32c"])
def test_autotune_keeps_acceleration_bound_dt(arm_name):
    """The tuner re-runs the momentum sweep, which rewrites the accelerations
    that the next step's critical step reads.  A scene whose dt is set by the
    |a| term, tuned in the middle of its run, ends bit-identical to the same
    scene untuned."""
    from sph07_amd.native import Scene, ARM_FP64, ARM_FP32C
    arm = ARM_FP64 if arm_name == "fp64" else ARM_FP32C
    deck = _acceleration_bound_channel()
    print(f"autotune: acceleration-bound channel {deck.np} particles, arm {arm_name}", flush=True)
    tuned, plain = Scene(deck, arm), Scene(deck, arm)
    plain.autotune(0)
    assert tuned.Solution(5) == plain.Solution(5) == 5
    # the acceleration term binds: critts below the sound and viscous limits
    h = deck.materials[0].h
    g = plain.globals()
    assert g["critts"] < 0.25 * h / 1.0e-3
    assert g["critts"] < 0.125 * h * h / deck.materials[0].av_l
    tuned.autotune(3)
    assert tuned.Solution(15) == 15
    assert plain.Solution(15) == 15
    _same(tuned, plain, f"{arm_name} tuned vs untuned")


def test_fp64_launch_shapes_same_bits():
    """The FP64 sweeps' records in flight per lane (2, 4, 8) change the
    schedule, not the arithmetic."""
    from sph07_amd.native import Scene
    deck = scenes.dam_break(nx_fluid=24, ny=40, nz_fluid=20, nx_tank=60, nz_tank=30, maxnbr=100)
    print(f"fp64 launch shapes: dam break {deck.np} particles", flush=True)
    ref = Scene(deck)
    ref.set_launch_shape()
    assert ref.Solution(20) == 20
    for gs, gm in ((2, 8), (8, 4), (4, 2)):
        other = Scene(deck)
        other.set_launch_shape(strain_batch=gs, momentum_batch=gm)
        assert other.Solution(20) == 20
        _same(ref, other, f"fp64 strain {gs} momentum {gm}")
        other.close()


def test_acceleration_bound_dt_matches_oracle():
    """The critical step's |a| term from the |a|^2 kept per particle (written
    with the velocity update and moved by the sort, while a itself stays
    behind in a fused step) against the oracle, which sums a of every
    particle each step: 40 fused steps (graph and direct launches) of a
    channel whose dt the |a| term sets, bit-identical, dt included."""
    from oracle import oracle as orc
    from sph07_amd.native import Scene
    deck = _acceleration_bound_channel()
    print(f"acceleration-bound dt vs oracle: {deck.np} particles", flush=True)
    o, g = orc.Oracle(deck), Scene(deck)
    assert o.run(40) == g.Solution(40) == 40
    h = deck.materials[0].h
    go, gg = o.globals(), g.globals()
    assert go["critts"] < 0.25 * h / 1.0e-3 and go["critts"] < 0.125 * h * h / deck.materials[0].av_l
    for k in ("dt", "dtold", "ptime", "critts"):
        assert gg[k] == go[k], (k, gg[k], go[k])
    from sph07_amd import native as nv
    for name, fg, fo in (("x", nv.F_X, orc.F_X), ("v", nv.F_V, orc.F_V), ("a", nv.F_A, orc.F_A),
                         ("rho", nv.F_RHO, orc.F_RHO)):
        assert np.array_equal(g.field(fg), o.field(fo)), name
