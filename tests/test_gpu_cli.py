"""The command-line driver (the reference's main.cpp on the device path),
built against the reference's class names in include/sph07_amd.hpp: the
fused loop (MSolution::Solution) and the same loop one phase call at a time
(MSolution.cpp:250-319) end on the same step, time and dt, bit for bit, and
agree with the library driven through ctypes and with the oracle."""
import os
import re
import subprocess

import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sph07_amd", "sph07_run")
DECK = os.path.join(ROOT, "tests", "golden", "kuglica.mcm")


def run_cli(mode, steps):
    print(f"cli: {os.path.basename(DECK)} {mode} {steps}", flush=True)
    r = subprocess.run([EXE, DECK, mode, str(steps)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    m = re.search(r"final: timestep (\d+) ptime (\S+) dt (\S+) grid (\d+) (\d+) (\d+)", r.stdout)
    assert m, r.stdout[-2000:]
    return int(m.group(1)), float(m.group(2)), float(m.group(3)), tuple(int(m.group(k)) for k in (4, 5, 6))


def test_fused_and_phase_by_phase_agree():
    from sph07_amd.native import Scene, read_mcm
    steps = 25
    fused = run_cli("fp64", steps)
    phased = run_cli("phases", steps)
    assert fused == phased, (fused, phased)
    assert fused[0] == steps + 1
    deck = read_mcm(DECK)
    g = Scene(deck)
    assert g.Solution(steps) == steps
    gl = g.globals()
    assert (gl["timestep"], gl["ptime"], gl["dt"]) == fused[:3]
    assert tuple(g.grid_info()["gridlim"]) == fused[3]
    o = orc.Oracle(deck)
    o.run(steps)
    og = o.globals()
    assert abs(og["ptime"] - fused[1]) <= 1e-10 * abs(og["ptime"])
    assert abs(og["dt"] - fused[2]) <= 1e-10 * abs(og["dt"])
    g.close()
