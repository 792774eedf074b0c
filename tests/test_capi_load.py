"""The C ABI library loads, exports every entry point declared in
include/sph07_amd.h, and its struct layouts match the Python deck packing.
No compute call: runs on a machine without a GPU."""
import ctypes
import os

import pytest

from sph07_amd import native
from sph07_amd.deck import CDeck, CMaterial


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(native.LIB_PATH):
        from sph07_amd import build
        build.build()
    return ctypes.CDLL(native.LIB_PATH)


def test_exports_every_declared_symbol(lib):
    names = native.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts(lib):
    assert lib.sph_desc_size() == ctypes.sizeof(CDeck)
    assert lib.sph_material_desc_size() == ctypes.sizeof(CMaterial)


def test_oracle_layout_matches():
    from oracle import oracle as orc
    L = orc.lib()
    assert L.orc_deck_size() == ctypes.sizeof(CDeck)
    assert L.orc_material_size() == ctypes.sizeof(CMaterial)


def test_no_cpu_fallback_in_product():
    """The product package never imports, links or loads the oracle."""
    import re
    pkg = os.path.dirname(native.__file__)
    pat = re.compile(r"(import\s+oracle|from\s+oracle|sph_oracle|libsph_oracle|orc_)")
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                txt = open(os.path.join(root, f)).read()
                assert not pat.search(txt), f


def test_cpp_surface_header_compiles(tmp_path):
    """include/sph07_amd.hpp (the reference's class names over the C ABI)
    compiles as C++17 and links against the library; nothing is run."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "use_surface.cpp"
    src.write_text(
        '#include "sph07_amd.hpp"\n'
        "int main() {\n"
        "  sph07_amd::MSimulationInit init;\n"
        "  sph07_amd::MSolution s(init.Context(), 2);\n"
        "  sph07_amd::MKernelBSpline *k = s.GenerateKernel(1.0);\n"
        "  int b[3] = {0, 0, 0}; int32_t lim[2][3];\n"
        "  s.NeighbourVars().GetLoopLimits(2, b, lim);\n"
        "  delete k;\n"
        "  return 0;\n"
        "}\n")
    exe = tmp_path / "use_surface"
    r = subprocess.run([gxx, "-std=c++17", "-I", os.path.join(root, "include"), str(src), "-o", str(exe),
                        "-L", os.path.dirname(native.LIB_PATH), "-lsph07_amd"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
