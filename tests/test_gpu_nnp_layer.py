"""Predictions through the header-compatible nnp layer on the GPU: the
layer's Prediction (setup, readStructureFromFile, predict) on the
reference's example NNPs against the reference's own goldens
(test/cpp/ExamplePrediction.h: energy 1e-10 relative, forces 1e-9 eV/A as
for the C ABI), and, where it was built, the reference's nnp-predict
application compiled unchanged against the layer, whose energy.out and
nnforces.out must hold the same numbers."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "n2p2_amd", "nnp", "bin")
E_REL = 1e-10
F_ABS = 1e-9


def say(*a):
    print("[gpu-nnp]", *a, flush=True)
    sys.stdout.flush()


def _golden(name):
    return cases.load_json("example_prediction.json")["examples"][name]


@pytest.mark.parametrize("name", cases.EXAMPLE_NAMES)
def test_prediction_through_layer(name):
    exe = os.path.join(BIN, "predict_example")
    assert os.path.exists(exe), "predict_example not built"
    say("Prediction via the nnp layer:", name)
    out = subprocess.run([exe, os.path.join(cases.EXAMPLES, name)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = out.stdout.splitlines()
    e = float(lines[0].split()[1])
    f = np.array([[float(x) for x in l.split()[2:5]] for l in lines[1:]])
    g = _golden(name)
    say("  energy %.12f vs %.12f" % (e, g["energy"]))
    assert abs(e - g["energy"]) <= E_REL * abs(g["energy"])
    assert np.abs(f - np.array(g["forces"])).max() <= F_ABS


@pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "nnp-predict")),
                    reason="the reference's nnp-predict was not built "
                           "against the layer (needs the reference checkout "
                           "at build time)")
@pytest.mark.parametrize("name", cases.EXAMPLE_NAMES)
def test_reference_nnp_predict_on_layer(tmp_path, name):
    say("reference nnp-predict on the layer:", name)
    src = os.path.join(cases.EXAMPLES, name)
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    out = subprocess.run([os.path.join(BIN, "nnp-predict"), "0"],
                         cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    g = _golden(name)
    rows = [l.split() for l in open(tmp_path / "energy.out")
            if not l.startswith("#")]
    e = float(rows[0][3])
    assert abs(e - g["energy"]) <= E_REL * abs(g["energy"])
    rows = [l.split() for l in open(tmp_path / "nnforces.out")
            if not l.startswith("#")]
    f = np.array([[float(x) for x in r[5:8]] for r in rows])
    assert np.abs(f - np.array(g["forces"])).max() <= F_ABS
    assert os.path.exists(tmp_path / "output.data")


def test_dataset_batch_through_layer(tmp_path):
    """Mode::evaluateNNPBatch on the 20 water structures of the reference's
    example data set with the example H2O_RPBE-D3 potential: every
    structure's energy and forces equal Mode::evaluateNNP of the structure
    alone (1e-12 relative energy, 1e-12 force units)."""
    src = os.path.join(cases.EXAMPLES, "H2O_RPBE-D3")
    for f in ("input.nn", "scaling.data", "weights.001.data",
              "weights.008.data"):
        shutil.copy(os.path.join(src, f), tmp_path)
    shutil.copy(os.path.join(cases.GOLDEN, "nnp_scaling_H2O", "input.data"),
                tmp_path)
    out = subprocess.run([os.path.join(BIN, "dataset_batch"), str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:], flush=True)
    assert out.returncode == 0
    rows = [l.split() for l in out.stdout.splitlines()
            if l.startswith("structure")]
    assert len(rows) == 20
    for w in rows:
        eb, es, df = float(w[3]), float(w[4]), float(w[5])
        assert abs(eb - es) <= 1e-12 * abs(es), w
        assert df <= 1e-12, w
