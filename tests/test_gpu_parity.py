"""Parity of the HIP FP64 arm against the CPU FP64 oracle, through the C ABI.

Scenes are small versions of every workload (periodic channel, 3D dam break
with the Monaghan term, rigid cylinder, lid-driven cavity, CSF droplet).
Bars (north-star): neighbour sets identical as sets -- and here also in
order; density, acceleration, velocity and position within 1e-10 relative
after N steps.  Scenes without a rigid body are expected bit-identical,
because the device sums each particle's neighbours in the reference order
with the reference's operation order and no FMA; the rigid-body reduction
uses a fixed tree order, so those scenes are held to the 1e-10 bar only.
"""
import numpy as np
import pytest

from sph07_amd import scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def small_scenes():
    return {
        "poiseuille": lambda: scenes.poiseuille(nx=40, ny=40),
        "dam_break": lambda: scenes.dam_break(nx_fluid=10, ny=8, nz_fluid=10, nx_tank=20,
                                              nz_tank=14, maxnbr=100),
        "cylinder": lambda: scenes.cylinder(nx=60, ny=100, dx=2e-4, radius=2e-3,
                                            centre=(0.006, 0.014)),
        "cavity": lambda: scenes.cavity(n=30),
        "droplet": lambda: scenes.droplet(n_box=40),
    }


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def gpu_scene(deck):
    from sph07_amd.native import Scene
    return Scene(deck)


FIELDS = [("x", orc.F_X), ("v", orc.F_V), ("a", orc.F_A), ("rho", orc.F_RHO), ("p", orc.F_P)]


def compare_state(g, o, rtol, exact, what):
    for name, f in FIELDS:
        a, b = g.field(f), o.field(f)
        if exact:
            assert np.array_equal(a, b), f"{what}: {name} differs (rel {rel(a, b):.3e})"
        else:
            assert rel(a, b) <= rtol, f"{what}: {name} rel {rel(a, b):.3e}"


@pytest.mark.parametrize("name", list(small_scenes()))
def test_initial_state(name):
    deck = small_scenes()[name]()
    o = orc.Oracle(deck)
    g = gpu_scene(deck)
    go, gg = o.globals(), g.globals()
    assert gg["dt"] == go["dt"], (gg["dt"], go["dt"])
    nn_o, c_o = o.neighbours()
    nn_g, c_g = g.neighbours()
    assert np.array_equal(nn_o, nn_g)
    for i in range(deck.np):
        assert np.array_equal(c_o[i, :nn_o[i]], c_g[i, :nn_g[i]]), f"particle {i} list differs"


@pytest.mark.parametrize("name", list(small_scenes()))
def test_phases_step1(name):
    """Every MSolution phase of the first two iterations, one at a time."""
    deck = small_scenes()[name]()
    exact = not deck.rigid_body
    o = orc.Oracle(deck)
    g = gpu_scene(deck)
    checks = {
        orc.PH_NEIGHBOURS: [],
        orc.PH_STRAIN: [("trd", orc.F_TRD), ("x", orc.F_X)],
        orc.PH_NORMALS: [("kappa", orc.F_KAPPA), ("normal", orc.F_NORMAL)] if deck.surface_tension else [],
        orc.PH_CONSTITUTIVE: [("rho", orc.F_RHO), ("p", orc.F_P)],
        orc.PH_TIMESTEP: [],
        orc.PH_MOMENTUM: [("a", orc.F_A)],
        orc.PH_BOUNDCOND: [("a", orc.F_A)],
        orc.PH_VELOCITY: [("v", orc.F_V)],
        orc.PH_MOVE: [("x", orc.F_X)],
    }
    for step in range(2):
        for ph in range(orc.NPHASES):
            o.phase(ph)
            g.phase(ph)
            if ph == orc.PH_NEIGHBOURS:
                nn_o, c_o = o.neighbours()
                nn_g, c_g = g.neighbours()
                assert np.array_equal(nn_o, nn_g), f"step {step}: neighbour counts differ"
                for i in range(deck.np):
                    assert np.array_equal(c_o[i, :nn_o[i]], c_g[i, :nn_g[i]])
            for fname, f in checks[ph]:
                a, b = g.field(f), o.field(f)
                if exact:
                    assert np.array_equal(a, b), f"step {step} phase {ph} {fname}: rel {rel(a, b):.3e}"
                else:
                    assert rel(a, b) <= RTOL, f"step {step} phase {ph} {fname}: rel {rel(a, b):.3e}"
        go, gg = o.globals(), g.globals()
        for k in ("dt", "dtold", "ptime", "critts"):
            assert gg[k] == go[k] or abs(gg[k] - go[k]) <= RTOL * abs(go[k]), (step, k, gg[k], go[k])
        assert gg["timestep"] == go["timestep"]


@pytest.mark.parametrize("name,steps", [("poiseuille", 200), ("dam_break", 60), ("cylinder", 60),
                                        ("cavity", 200), ("droplet", 100)])
def test_fused_steps(name, steps):
    """The fused device-resident loop (sph_solution) against the oracle."""
    deck = small_scenes()[name]()
    o = orc.Oracle(deck)
    g = gpu_scene(deck)
    n_o = o.run(steps)
    n_g = g.Solution(steps)
    assert n_o == n_g == steps
    compare_state(g, o, RTOL, exact=not deck.rigid_body, what=f"{name} after {steps} steps")
    go, gg = o.globals(), g.globals()
    for k in ("dt", "ptime", "cmm_min0", "cmm_max0", "cmm_min1", "cmm_max1"):
        assert abs(gg[k] - go[k]) <= RTOL * max(abs(go[k]), 1e-300), (k, gg[k], go[k])
    if deck.rigid_body:
        # scales: |R| for the centre, |V| for velocities, |V|/radius for the
        # spin (zero up to rounding for this symmetric body)
        R = np.hypot(go["rbR0"], go["rbR1"])
        V = max(np.hypot(go["rbV0"], go["rbV1"]), 1e-300)
        scale = {"rbR0": R, "rbR1": R, "rbV0": V, "rbV1": V, "rbW2": V / deck.meta["radius"]}
        for k, sc in scale.items():
            assert abs(gg[k] - go[k]) <= 1e-9 * sc, (k, gg[k], go[k])


@pytest.mark.parametrize("hfac,steps", [(4.5, 30), (6.0, 20)])
def test_dense_cells(hfac, steps):
    """Large h/dx: 2h cells of 81 (h = 4.5 dx) and 144 (h = 6 dx) particles.
    The first has more targets per cell than a wave (two target groups per
    cell); the second has stencils of ~1300 candidates, beyond the cell
    kernel's staging space, so those cells take the per-particle walk.  Both
    stay bit-identical to the oracle, lists included."""
    deck = scenes.cavity(n=24, hfac=hfac, maxnbr=700)
    o = orc.Oracle(deck)
    g = gpu_scene(deck)
    nn_o, ids_o = o.neighbours()
    nn_g, ids_g = g.neighbours()
    assert np.array_equal(nn_o, nn_g)
    assert nn_g.max() > 64
    for i in range(len(nn_o)):
        assert np.array_equal(ids_o[i, :nn_o[i]], ids_g[i, :nn_g[i]]), i
    assert o.run(steps) == g.Solution(steps) == steps
    compare_state(g, o, RTOL, exact=True, what=f"dense cells h={hfac}dx after {steps} steps")


def _two_material_dam(tcrit_opt):
    """The small dam break with the fluid's far half (y) given a second
    material whose h is 1.25 times the first: pairs of unlike particles use
    the mean h, so the search has three cut-offs."""
    import dataclasses
    deck = scenes.dam_break(nx_fluid=10, ny=8, nz_fluid=10, nx_tank=20, nz_tank=14, maxnbr=200)
    m1 = deck.materials[0]
    dx3 = deck.meta["dx"] ** 3
    second = np.zeros(deck.np, bool)
    second[:deck.n_fluid] = deck.x[:deck.n_fluid, 1] > 0.5 * deck.x[:deck.n_fluid, 1].max()
    n2 = int(second.sum())
    m2 = dataclasses.replace(m1, h=1.25 * m1.h, mass=m1.rho * dx3 * n2)
    deck.materials = [dataclasses.replace(m1, mass=m1.rho * dx3 * (deck.np - n2)), m2]
    deck.mat_id = np.where(second, 2, 1).astype(np.int32)
    deck.tcrit_opt = tcrit_opt
    return deck


def _with_tcrit(deck, tcrit_opt):
    deck.tcrit_opt = tcrit_opt
    return deck


CELL_KERNEL_CASES = {
    # critical step from the smallest neighbour distance (tcrit option 2)
    "dam_break_mindist": lambda: _with_tcrit(small_scenes()["dam_break"](), 2),
    "cylinder_mindist": lambda: _with_tcrit(small_scenes()["cylinder"](), 2),
    # per-pair cut-offs
    "dam_break_two_h": lambda: _two_material_dam(3),
    "dam_break_two_h_mindist": lambda: _two_material_dam(2),
}


@pytest.mark.parametrize("name", list(CELL_KERNEL_CASES))
def test_cell_kernel_variants(name):
    """The neighbour search's template variants (one or several cut-offs, with
    or without the minimum distance) against the oracle: lists identical,
    then 30 fused steps (bit-identical without a rigid body, dt included)."""
    deck = CELL_KERNEL_CASES[name]()
    o = orc.Oracle(deck)
    g = gpu_scene(deck)
    nn_o, ids_o = o.neighbours()
    nn_g, ids_g = g.neighbours()
    assert np.array_equal(nn_o, nn_g)
    for i in range(deck.np):
        assert np.array_equal(ids_o[i, :nn_o[i]], ids_g[i, :nn_g[i]]), i
    assert g.globals()["dt"] == o.globals()["dt"]
    assert o.run(30) == g.Solution(30) == 30
    compare_state(g, o, RTOL, exact=not deck.rigid_body, what=name)
    go, gg = o.globals(), g.globals()
    for k in ("dt", "critts"):
        # critts: with tcrit option 2 the smallest mindist / (c + |v|)
        assert abs(gg[k] - go[k]) <= RTOL * go[k], (k, gg[k], go[k])
        if not deck.rigid_body:
            assert gg[k] == go[k], (k, gg[k], go[k])


def test_fp32c_two_materials():
    """FP32c arm with two materials of different h (compact momentum records
    with a mass per material) against the FP64 arm."""
    from sph07_amd.native import Scene, ARM_FP32C, F_V
    deck = _two_material_dam(3)
    g64, g32 = Scene(deck), Scene(deck, ARM_FP32C)
    assert g64.Solution(60) == g32.Solution(60) == 60
    moving = deck.dispbc == 0
    v64, v32 = g64.field(F_V)[moving], g32.field(F_V)[moving]
    err = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    print(f"two materials: fp32c rel L2 velocity error {err:.3e}")
    assert err <= 1e-5, err


@pytest.mark.parametrize("hfac,steps", [(3.0, 40), (4.5, 30), (6.0, 20)])
def test_fp32c_dense_cells(hfac, steps):
    """The FP32c sweeps over long lists (h = 3 dx: ~110 neighbours), lists
    from the cell kernel's two target groups per cell (h = 4.5 dx) and from
    the per-particle walk of oversized stencils (h = 6 dx): against the FP64
    arm, neighbour counts shared."""
    from sph07_amd.native import Scene, ARM_FP32C, F_V, F_RHO
    deck = scenes.cavity(n=24, hfac=hfac, maxnbr=700)
    g64, g32 = Scene(deck), Scene(deck, ARM_FP32C)
    assert g64.Solution(steps) == g32.Solution(steps) == steps
    moving = deck.dispbc == 0
    v64, v32 = g64.field(F_V)[moving], g32.field(F_V)[moving]
    err = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    print(f"h = {hfac} dx: fp32c rel L2 velocity error {err:.3e}")
    assert err <= 1e-5, err
    r64, r32 = g64.field(F_RHO), g32.field(F_RHO)
    assert np.abs(r32 - r64).max() / np.abs(r64).max() <= 1e-6
    nn64, _ = g64.neighbours()
    nn32, _ = g32.neighbours()
    assert np.array_equal(nn64, nn32)


def test_neighbour_overflow_is_reported():
    """maxnbr too small: the reference's 'Neighbour list exceeded maximum.'"""
    from sph07_amd.native import SphError
    deck = scenes.poiseuille(nx=20, ny=20, maxnbr=8)
    with pytest.raises(SphError) as e:
        gpu_scene(deck)
    assert e.value.code == 2 and "exceeded maximum" in str(e.value)


@pytest.mark.parametrize("name,steps", [("poiseuille", 200), ("dam_break", 60), ("cylinder", 60),
                                        ("cavity", 200), ("droplet", 100)])
def test_fp32c_arm_error(name, steps):
    """FP32-compensated pair sums: relative L2 velocity error of the moving
    particles <= 1e-5 against the FP64 arm (which matches the oracle above);
    neighbour sets identical."""
    from sph07_amd.native import Scene, ARM_FP32C, F_V, F_RHO
    deck = small_scenes()[name]()
    g64 = Scene(deck)
    g32 = Scene(deck, ARM_FP32C)
    assert g64.Solution(steps) == g32.Solution(steps) == steps
    moving = deck.dispbc == 0
    v64, v32 = g64.field(F_V)[moving], g32.field(F_V)[moving]
    err = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    print(f"{name}: fp32c rel L2 velocity error {err:.3e}")
    assert err <= 1e-5, err
    r64, r32 = g64.field(F_RHO), g32.field(F_RHO)
    assert np.abs(r32 - r64).max() / np.abs(r64).max() <= 1e-6
    nn64, c64 = g64.neighbours()
    nn32, c32 = g32.neighbours()
    assert np.array_equal(nn64, nn32)


def test_fp32c_arm_error_varied_masses():
    """Per-particle masses (massopt 1) that differ within a material: the
    FP32c momentum sweep uses its 40-byte records (mass and p/rho^2 stored)."""
    from sph07_amd.native import Scene, ARM_FP32C, F_V
    deck = scenes.dam_break(nx_fluid=10, ny=8, nz_fluid=10, nx_tank=20, nz_tank=14, maxnbr=100)
    m0 = deck.materials[0].mass / deck.np
    k = np.arange(deck.np)
    deck.node_mass = m0 * (1.0 + 0.01 * np.sin(0.37 * k))
    deck.massopt = 1
    g64, g32 = Scene(deck), Scene(deck, ARM_FP32C)
    assert g64.Solution(60) == g32.Solution(60) == 60
    moving = deck.dispbc == 0
    v64, v32 = g64.field(F_V)[moving], g32.field(F_V)[moving]
    err = np.linalg.norm(v32 - v64) / np.linalg.norm(v64)
    print(f"varied masses: fp32c rel L2 velocity error {err:.3e}")
    assert err <= 1e-5, err


def golden_deck(name, tmp_path):
    """A reference deck from tests/golden (the larger ones gzip-compressed),
    read by the library's .mcm reader."""
    import gzip
    import os
    from sph07_amd.native import read_mcm
    path = os.path.join(os.path.dirname(__file__), "golden", name)
    if name.endswith(".gz"):
        out = tmp_path / name[:-3]
        with gzip.open(path, "rb") as f:
            out.write_bytes(f.read())
        path = str(out)
    return read_mcm(path)


# the reference's own input decks that lie on the path: rigid body (rigidbody,
# decuzzi: with periodic x), CSF (kuglica; rbcell: with periodic x; rbcell3d:
# 3D), the Tait EOS 28 (stenosis, cylinders), periodic flow past obstacles
# (flowpast)
REFERENCE_DECKS = [("rigidbody.mcm", 100), ("kuglica.mcm", 100), ("stenosis.mcm", 100), ("rbcell.mcm", 60),
                   ("cylinders.mcm.gz", 60), ("flowpast.mcm.gz", 60), ("decuzzi.mcm.gz", 40),
                   ("rbcell3d.mcm.gz", 20)]


@pytest.mark.parametrize("deckfile,steps", REFERENCE_DECKS)
def test_reference_decks(deckfile, steps, tmp_path):
    """The reference's own input decks: device FP64 arm against the oracle
    after `steps` iterations.  Bit-identical unless the deck has a rigid body
    (fixed-tree sums) or the Tait EOS 28, whose pow() is correctly rounded
    neither by the host's libm nor on the device (one-ulp differences on some
    inputs); those are held to the 1e-10 bar."""
    deck = golden_deck(deckfile, tmp_path)
    o = orc.Oracle(deck)
    g = gpu_scene(deck)
    nn_o, ids_o = o.neighbours()
    nn_g, ids_g = g.neighbours()
    assert np.array_equal(nn_o, nn_g)
    assert o.run(steps) == g.Solution(steps) == steps
    tait = any(m.eos == 28 for m in deck.materials)
    compare_state(g, o, RTOL, exact=not deck.rigid_body and not tait, what=deckfile)


@pytest.mark.parametrize("arm_name", ["fp64", "fp32c"])
def test_autotune_changes_timing_not_results(arm_name):
    """A scene tuned in the middle of a run ends bit-identical to the same
    scene with the tuning turned off."""
    from sph07_amd.native import Scene, ARM_FP64, ARM_FP32C, F_X, F_V, F_RHO
    arm = ARM_FP64 if arm_name == "fp64" else ARM_FP32C
    deck = scenes.dam_break(nx_fluid=24, ny=40, nz_fluid=20, nx_tank=60, nz_tank=30, maxnbr=100)
    tuned, plain = Scene(deck, arm), Scene(deck, arm)
    plain.autotune(0)
    assert tuned.Solution(5) == 5
    shape = tuned.autotune(3)
    assert tuned.Solution(15) == 15
    assert plain.Solution(20) == 20
    assert shape["cell_slots"] in (16, 32, 64, 128)
    for f in (F_X, F_V, F_RHO):
        assert np.array_equal(tuned.field(f), plain.field(f))


@pytest.mark.parametrize("arm_name", ["fp64", "fp32c"])
@pytest.mark.parametrize("name", ["dam_break", "poiseuille", "droplet"])
def test_cell_order_changes_timing_not_results(arm_name, name):
    """The cells stored in row-major instead of Morton order: the same bits
    after 30 fused steps (the neighbour lists follow the traversal order, not
    the storage order)."""
    from sph07_amd.native import Scene, ARM_FP64, ARM_FP32C, F_X, F_V, F_RHO
    arm = ARM_FP64 if arm_name == "fp64" else ARM_FP32C
    deck = small_scenes()[name]()
    morton, rowmajor = Scene(deck, arm), Scene(deck, arm)
    morton.set_launch_shape(cell_order="morton")
    rowmajor.set_launch_shape(cell_order="row-major", cell_slots=32)
    assert morton.Solution(30) == rowmajor.Solution(30) == 30
    for f in (F_X, F_V, F_RHO):
        assert np.array_equal(morton.field(f), rowmajor.field(f)), f
    nm, cm = morton.neighbours()
    nr, cr = rowmajor.neighbours()
    assert np.array_equal(nm, nr) and np.array_equal(cm, cr)


@pytest.mark.parametrize("strain_g,momentum_g", [(4, 2), (8, 8), (4, 4)])
def test_fp32c_launch_shapes_same_bits(strain_g, momentum_g):
    """The FP32c sweeps' records in flight per lane change the schedule, not
    the arithmetic: every batch size gives the bits of strain 8 / momentum 4
    (the sweeps fuse multiply-adds per source expression only)."""
    from sph07_amd.native import Scene, ARM_FP32C, F_X, F_V, F_RHO
    deck = scenes.dam_break(nx_fluid=24, ny=40, nz_fluid=20, nx_tank=60, nz_tank=30, maxnbr=100)
    ref, other = Scene(deck, ARM_FP32C), Scene(deck, ARM_FP32C)
    ref.set_launch_shape(strain_batch=8, momentum_batch=4)
    other.set_launch_shape(strain_batch=strain_g, momentum_batch=momentum_g)
    assert ref.Solution(20) == other.Solution(20) == 20
    for f in (F_X, F_V, F_RHO):
        assert np.array_equal(ref.field(f), other.field(f)), f


@pytest.mark.parametrize("name", ["dam_break", "poiseuille"])
def test_walk_shape_same_bits(name):
    """Slots per wave 0 (the per-particle neighbour walk, an autotuner
    candidate) builds the same lists in the same order as the cell kernel:
    20 fused steps give the same bits."""
    from sph07_amd.native import Scene
    deck = (scenes.dam_break(nx_fluid=10, ny=8, nz_fluid=10, nx_tank=20, nz_tank=14, maxnbr=100)
            if name == "dam_break" else scenes.poiseuille(nx=30, ny=30))
    print(f"walk shape: {name}, {deck.np} particles", flush=True)
    cell, walk = Scene(deck), Scene(deck)
    cell.set_launch_shape(cell_slots=32)
    walk.set_launch_shape(cell_slots=0)
    assert cell.Solution(20) == walk.Solution(20) == 20
    nn_c, ids_c = cell.neighbours()
    nn_w, ids_w = walk.neighbours()
    assert np.array_equal(nn_c, nn_w) and np.array_equal(ids_c, ids_w)
    for f in (0, 1, 2, 3):
        assert np.array_equal(cell.field(f), walk.field(f)), f
    # the phase entry point searches again on the final state (the walk's 2D
    # positions refreshed from x)
    cell.LinkedListNeighbours()
    walk.LinkedListNeighbours()
    nn_c, ids_c = cell.neighbours()
    nn_w, ids_w = walk.neighbours()
    assert np.array_equal(nn_c, nn_w) and np.array_equal(ids_c, ids_w)
