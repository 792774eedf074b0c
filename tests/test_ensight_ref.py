"""Known answers for the EnSight writer model that checks the library's state
output (tests/ensight_ref.py): seven particles of two materials, the lines
written out by hand from the reference's print formats (MOutput.cpp): nodes
in pairs per line (:353-364), per-material values six per line, broken on
the particle number in 3D (:618-633) and on the count in 1D/2D (:459-470),
and the case file's time list (:147-154)."""
import numpy as np

from ensight_ref import EnsightModel

MAT = [1, 2, 1, 1, 2, 1, 1]


def fields(n=7):
    i = np.arange(1, n + 1, dtype=float)
    x = np.stack([i * 1e-3, np.zeros(n), -i * 1e-3], axis=1)
    z = np.zeros(n)
    return {"x": x, "v": np.zeros((n, 3)), "a": np.zeros((n, 3)), "mass": np.full(n, 0.5), "boundary": z,
            "nnbr": i, "rho": 1000.0 + i, "efps": z, "energy": z, "c": np.full(n, 1.5), "temper": z,
            "sigma": np.stack([-i, -2 * i, 0.25 * i], axis=1)}


def test_density_breaks_on_the_particle_number_in_3d():
    m = EnsightModel("b", 3, MAT, 2, np.zeros((7, 3)))
    m.state(fields(), 0.0)
    assert m.files["_rho.ens"] == (
        "BEGIN TIME STEP\n"
        "density\n"
        "part        1\n"
        "point\n"
        "  1.0010e+03  1.0030e+03  1.0040e+03  1.0060e+03\n"
        "  1.0070e+03\n"
        "part        2\n"
        "point\n"
        "  1.0020e+03  1.0050e+03\n"
        "END TIME STEP\n")


def test_density_breaks_on_the_count_in_2d():
    m = EnsightModel("b", 2, MAT, 2, np.zeros((7, 3)))
    m.state(fields(), 0.0)
    assert m.files["_rho.ens"] == (
        "BEGIN TIME STEP\n"
        "density\n"
        "part        1\n"
        "point\n"
        "  1.0010e+03  1.0030e+03  1.0040e+03  1.0060e+03  1.0070e+03\n"
        "part        2\n"
        "point\n"
        "  1.0020e+03  1.0050e+03\n"
        "END TIME STEP\n")


def test_displacement_pairs_and_stress_rows():
    m = EnsightModel("b", 3, MAT, 2, np.zeros((7, 3)))
    m.state(fields(), 0.0)
    assert m.files["_dis.ens"] == (
        "BEGIN TIME STEP\n"
        "displacement\n"
        "   1.000e-03   0.000e+00  -1.000e-03   2.000e-03   0.000e+00  -2.000e-03\n"
        "   3.000e-03   0.000e+00  -3.000e-03   4.000e-03   0.000e+00  -4.000e-03\n"
        "   5.000e-03   0.000e+00  -5.000e-03   6.000e-03   0.000e+00  -6.000e-03\n"
        "   7.000e-03   0.000e+00  -7.000e-03\n"
        "END TIME STEP\n")
    sig = m.files["_sig.ens"].split("\n")
    assert sig[2:5] == ["part        1", "point",
                        " -1.0000e+00 -2.0000e+00  0.0000e+00  2.5000e-01  0.0000e+00  0.0000e+00"]
    assert m.files["_nbr.ens"].split("\n")[2:4] == [
        "   1.000e+00   2.000e+00   3.000e+00   4.000e+00   5.000e+00   6.000e+00", "   7.000e+00"]


def test_case_time_list_and_appends():
    m = EnsightModel("b", 1, MAT, 2, np.zeros((7, 3)))
    for t in (0.0, 0.5, 1.25):
        m.state(fields(), t)
    case = m.files[".case"]
    assert ("TIME\n"
            "time set:                1\n"
            "number of steps:         3\n"
            "filename start number:   1\n"
            "filename increment:      1\n"
            "time values:   0.00000e+00\n"
            "               5.00000e-01\n"
            "               1.25000e+00\n"
            "\n"
            "FILE\n"
            "file set:                1\n"
            "filename index:          1\n"
            "number of steps:         3\n") in case
    assert case.startswith("FORMAT\ntype:   ensight\n\nGEOMETRY\nmodel:                                    "
                           "                     b.geo\n\nVARIABLES\n")
    assert m.files["_time.txt"] == " 0.000000e+00\n 5.000000e-01\n 1.250000e+00\n"
    assert m.files["_vel.ens"].count("BEGIN TIME STEP") == 3
    # 1D: the stress row holds sigma_xx only
    assert " -1.0000e+00  0.0000e+00  0.0000e+00  0.0000e+00  0.0000e+00  0.0000e+00" in m.files["_sig.ens"]
    # the geometry is written with the first state only
    assert m.files[".geo"].count("BEGIN TIME STEP") == 1
    assert m.files[".geo"].split("\n")[6:8] == ["7", "       1  1.00000e-03  0.00000e+00 -1.00000e-03"]
