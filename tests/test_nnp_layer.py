"""Host checks of the header-compatible nnp layer (n2p2_amd/nnp): the
input.data reader with its remap against the Python reader (itself pinned to
the oracle's neighbour lists), and that a caller of the reference compiles
unchanged against the layer's headers (the reference's nnp-predict
application, compiled from its source file where it lies; skipped where the
reference checkout is absent)."""
import os
import subprocess

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import io as nio
from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "n2p2_amd")
BIN = os.path.join(PKG, "nnp", "bin")
REF_PREDICT = "/root/reference/src/application/nnp-predict.cpp"


def _tool(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.fail("%s not built (make -C n2p2_amd)" % p)
    return p


@pytest.mark.parametrize("name", cases.EXAMPLE_NAMES)
def test_structure_reader_bitwise(name):
    d = os.path.join(cases.EXAMPLES, name)
    out = subprocess.run([_tool("read_structure"), os.path.join(d, "input.nn"),
                          os.path.join(d, "input.data")],
                         capture_output=True, text=True, check=True).stdout
    lines = out.splitlines()
    box = np.array([[float.fromhex(x) for x in l.split()[1:]]
                    for l in lines if l.startswith("box")])
    atoms = [l.split() for l in lines if l.startswith("atom ")]
    r = np.array([[float.fromhex(x) for x in a[3:6]] for a in atoms])
    el = np.array([int(a[2]) for a in atoms])
    m = engine.Model.load_dir(d)
    r0, el0, box0, _ = nio.read_structure(os.path.join(d, "input.data"),
                                          m.element_index)
    if box0 is None:            # open structure (DMABN_SCAN): no lattice
        assert not box.any()
    else:
        assert np.array_equal(box, box0)
    assert np.array_equal(el, el0)
    assert np.array_equal(r, r0)


def test_file_header_layout(tmp_path):
    """createFileHeader: column numbers and names right-aligned over columns
    whose widths are the given sizes (the first two narrower)."""
    src = tmp_path / "hdr.cpp"
    src.write_text(r'''
#include "utility.h"
#include <cstdio>
int main() {
    auto h = nnp::createFileHeader({"Energy comparison."}, {10, 10, 24},
                                   {"conf", "natoms", "Eref"},
                                   {"a", "b", "c"});
    for (auto const& l : h) std::fputs(l.c_str(), stdout);
}
''')
    exe = tmp_path / "hdr"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(PKG, "nnp", "include"),
                    str(src), "-o", str(exe), "-L", PKG, "-lnnp_amd", "-lhdnnp",
                    "-Wl,-rpath," + PKG], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True,
                           check=True).stdout.splitlines()
    assert lines[0] == "#" * 80 and lines[1] == "# Energy comparison."
    assert lines[3] == "# Col  Name   Description"
    assert lines[5] == "# 1    conf   a"
    widths = [8, 10, 24]
    num = "#" + "".join(" " + str(i + 1).rjust(w) for i, w in enumerate(widths))
    nam = "#" + "".join(" " + n.rjust(w)
                        for n, w in zip(["conf", "natoms", "Eref"], widths))
    assert lines[-3] == num and lines[-2] == nam
    assert lines[-1] == "#" * (10 + 10 + 24 + 2)


@pytest.mark.skipif(not os.path.exists(REF_PREDICT),
                    reason="reference checkout not present")
def test_reference_nnp_predict_compiles_unchanged(tmp_path):
    exe = tmp_path / "nnp-predict"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I",
                        os.path.join(PKG, "nnp", "include"), REF_PREDICT,
                        "-o", str(exe), "-L", PKG, "-lnnp_amd", "-lhdnnp",
                        "-Wl,-rpath," + PKG], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert exe.exists()
