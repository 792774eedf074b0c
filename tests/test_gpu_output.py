"""State output of the device path (the reference's MOutput, EnSight CASE
format, MOutput.cpp:39-720; the cylinder plot of MSolution.cpp:1638-1668; a
rigid trajectory line; a full-precision state dump).

No reference-produced file exists to compare against (the reference cannot be
built here), so the files are checked for the reference's structure -- the
file set, the time-step blocks, values per line, parts per material -- and
their values against the fields downloaded at the same moment, to the
precision the formats print (%12.3e: 4 digits, %12.4e: 5, %12.5e: 6).  The
dump is hexadecimal, so it must equal the downloads bit for bit."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "sph07_amd", "sph07_run")


def deck_copy(tmp_path, name="kuglica.mcm", stpltime=" 5.000E-04"):
    """The reference deck with a short time between state plots (card 3)."""
    src = open(os.path.join(GOLD, name)).read().split("\n")
    # card 3 is the 13th line of kuglica.mcm (title, comments, cards 1-2)
    k = next(i for i, ln in enumerate(src) if ln.startswith(" 0.500E+00    2"))
    src[k] = stpltime + src[k][10:]
    path = tmp_path / name
    path.write_text("\n".join(src))
    return str(path)


def blocks(path):
    """The time-step blocks of an EnSight variable file: lists of tokens."""
    txt = open(path).read()
    out = []
    for blk in re.findall(r"BEGIN TIME STEP\n(.*?)END TIME STEP\n", txt, re.S):
        out.append(blk.split("\n"))
    return out


def numbers(lines):
    vals = []
    for ln in lines:
        # fixed-width fields of 12 characters
        for k in range(0, len(ln), 12):
            f = ln[k:k + 12].strip()
            if f:
                vals.append(float(f))
    return np.array(vals)


def parts(lines):
    """Per-material sections of a block: {material: value lines}."""
    out, cur = {}, None
    for ln in lines:
        if ln.startswith("part "):
            cur = int(ln.split()[1])
            out[cur] = []
        elif cur is not None and ln and ln != "point":
            out[cur].append(ln)
    return out


def close(a, b, digits):
    a, b = np.asarray(a, float), np.asarray(b, float)
    tol = 0.6 * 10.0 ** (1 - digits) * np.abs(b) + 1e-300
    return np.all(np.abs(a - b) <= tol)


def test_ensight_files(tmp_path):
    from sph07_amd.native import (Scene, StateOutput, read_mcm, F_X, F_V, F_A, F_RHO, F_P, F_MASS,
                                  F_NNBR, F_C)
    path = deck_copy(tmp_path)
    deck = read_mcm(path)
    g = Scene(deck)
    base = str(tmp_path / "out")
    out = StateOutput(g, base, deck)
    x0 = g.field(F_X)
    snaps = []
    for steps in (0, 12, 7):
        if steps:
            assert g.Solution(steps) == steps
        out.state()
        snaps.append({f: g.field(f) for f in (F_X, F_V, F_A, F_RHO, F_P, F_MASS, F_NNBR, F_C)} |
                     {"t": g.globals()["ptime"]})
    out.cylinder_plot(base + ".drp")
    out.close()
    n = deck.np
    # case file: three states, the time values of MOutput::WriteCase
    case = open(base + ".case").read()
    assert "number of steps:         3" in case
    times = [float(t) for t in re.findall(r"^\s*(?:time values:)?\s*([-+.0-9eE]+)\s*$",
                                          case.split("TIME\n")[1].split("FILE")[0], re.M)]
    assert times[0] == 0.0 and abs(times[-1] - snaps[-1]["t"]) <= 1e-5 * snaps[-1]["t"]
    assert len(open(base + "_time.txt").read().split()) == 3
    # geometry: the first state's coordinates, one part per material
    geo = open(base + ".geo").read().split("\n")
    assert geo[6] == str(n)
    xyz = np.array([[float(v) for v in ln.split()[1:]] for ln in geo[7:7 + n]])
    assert close(xyz, x0, 6)
    assert geo.count("point") == len(deck.materials)
    # node variables: 2 nodes (6 values) per line, then the per-material ones
    for ext, f, digits in (("_vel.ens", F_V, 4), ("_acc.ens", F_A, 4)):
        bl = blocks(base + ext)
        assert len(bl) == 3
        for b, s in zip(bl, snaps):
            vals = numbers(b[1:])
            assert vals.size == 3 * n and close(vals.reshape(n, 3), s[f], digits), ext
    dis = blocks(base + "_dis.ens")
    for b, s in zip(dis, snaps):
        assert close(numbers(b[1:]).reshape(n, 3), s[F_X] - x0, 4) or \
            np.abs(numbers(b[1:]).reshape(n, 3) - (s[F_X] - x0)).max() < 1e-9
    for ext, f in (("_mas.ens", F_MASS), ("_nbr.ens", F_NNBR)):
        for b, s in zip(blocks(base + ext), snaps):
            assert all(len(ln) <= 72 for ln in b[1:])
            assert close(numbers(b[1:]), s[f], 4), ext
    mats = range(1, len(deck.materials) + 1)
    for ext, f in (("_rho.ens", F_RHO), ("_snd.ens", F_C)):
        for b, s in zip(blocks(base + ext), snaps):
            assert b[1] == "part        1" and b[2] == "point"
            pp = parts(b[1:])
            assert sorted(pp) == list(mats)
            for m in mats:
                assert close(numbers(pp[m]), s[f][deck.mat_id == m], 5), (ext, m)
    for b, s in zip(blocks(base + "_sig.ens"), snaps):
        pp = parts(b[1:])
        for m in mats:
            sig = numbers(pp[m]).reshape(-1, 6)
            pm = s[F_P][deck.mat_id == m]
            assert close(sig[:, 0], -pm, 5) and close(sig[:, 1], -pm, 5)
            assert np.all(sig[:, 2:] == 0.0)
    # the cylinder plot: x(0) and rho of particles 1 .. np-1
    drp = np.loadtxt(base + ".drp")
    assert drp.shape == (n - 1, 2)
    assert close(drp[:, 0], snaps[-1][F_X][:n - 1, 0], 6) and close(drp[:, 1], snaps[-1][F_RHO][:n - 1], 6)
    g.close()


def test_state_dump_is_exact(tmp_path):
    from sph07_amd.native import Scene, read_mcm, write_state_dump, read_state_dump, F_X, F_V, F_A, F_RHO, F_P
    deck = read_mcm(os.path.join(GOLD, "rigidbody.mcm"))
    g = Scene(deck)
    assert g.Solution(15) == 15
    p = str(tmp_path / "state.txt")
    write_state_dump(g, p)
    d = read_state_dump(p)
    gl = g.globals()
    assert d["np"] == deck.np and d["timestep"] == gl["timestep"]
    assert d["ptime"] == gl["ptime"] and d["dt"] == gl["dt"]
    for key, f in (("x", F_X), ("v", F_V), ("a", F_A), ("rho", F_RHO), ("p", F_P)):
        assert np.array_equal(d[key], g.field(f)), key
    g.close()


def test_rigid_trajectory_line(tmp_path):
    from sph07_amd.native import Scene, StateOutput, read_mcm
    deck = read_mcm(os.path.join(GOLD, "rigidbody.mcm"))
    g = Scene(deck)
    out = StateOutput(g, str(tmp_path / "rb"), deck)
    path = str(tmp_path / "rb.rbt")
    for _ in range(3):
        g.Solution(5)
        out.rigid(path)
    rows = np.loadtxt(path)
    assert rows.shape == (3, 10)
    gl = g.globals()
    assert abs(rows[-1, 0] - gl["ptime"]) <= 1e-5 * gl["ptime"]
    assert abs(rows[-1, 1] - gl["rbR0"]) <= 1e-5 * abs(gl["rbR0"])
    out.close()
    g.close()


def test_cli_output_and_dump(tmp_path):
    """sph07_run --output writes a state at t = 0 and after the step that
    passes each state-plot time; its final dump equals the library driven
    through ctypes for the same number of steps."""
    from sph07_amd.native import Scene, read_mcm, read_state_dump, F_X, F_V, F_RHO
    path = deck_copy(tmp_path)
    dump = str(tmp_path / "final.txt")
    steps = 30
    print(f"cli output: {path} {steps} steps", flush=True)
    r = subprocess.run([EXE, path, "fp64", str(steps), "--output", f"--dump={dump}"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    written = r.stdout.count("State plot written")
    assert written >= 4
    case = open(path + ".case").read()
    assert f"number of steps:     {written + 1:5d}" in case
    assert len(blocks(path + "_rho.ens")) == written + 1
    d = read_state_dump(dump)
    g = Scene(read_mcm(path))
    assert g.Solution(steps) == steps
    for key, f in (("x", F_X), ("v", F_V), ("rho", F_RHO)):
        assert np.array_equal(d[key], g.field(f)), key
    # state times: every written state's time passes the previous plot time
    t = [float(v) for v in open(path + "_time.txt").read().split()]
    assert t[0] == 0.0 and all(t[k] > 5e-4 * k for k in range(1, len(t)))
    g.close()
