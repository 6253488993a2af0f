"""The GPU trainer (n2p2_amd/train.py) on the reference's example data set
(tests/golden/nnp_scaling_H2O: 20 water structures with energies and
forces, its input.nn and scaling.data; the settings ask for the Kalman
filter, updater_type 1).

- Three epochs from random weights lower the training RMSE of energies and
  forces (the reference's own test only checks that nnp-train runs).
- The weights written after the last epoch (weights.%03zu.%06zu.out, the
  reference's file names and connection order) reload into a model that
  gives the same energies as the trained context.
"""
import os
import shutil
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd.train import Training

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "nnp_scaling_H2O")


def say(*a):
    print("[gpu-train-tool]", *a, flush=True)
    sys.stdout.flush()


def test_trainer_lowers_rmse_and_writes_weights(tmp_path):
    for f in ("input.nn", "input.data", "scaling.data"):
        shutil.copy(os.path.join(GOLD, f), tmp_path)
    t = Training(str(tmp_path))
    t.setup()
    t.read_data()
    curve = t.run(epochs=3, log=say)
    assert len(curve) == 4
    e0, e3 = curve[0], curve[-1]
    assert e3[1] < e0[1] and e3[3] < e0[3]
    lc = [l for l in open(tmp_path / "learning-curve.out")
          if not l.startswith("#")]
    assert len(lc) == 4
    d2 = tmp_path / "reload"
    os.makedirs(d2)
    for f in ("input.nn", "scaling.data"):
        shutil.copy(tmp_path / f, d2)
    for z in t.zs:
        shutil.copy(tmp_path / ("weights.%03d.%06d.out" % (z, 3)),
                    d2 / ("weights.%03d.data" % z))
    m = engine.Model.load_dir(str(d2))
    c = engine.Context(m)
    for k in t.test + t.train_set[:3]:
        r, elem, box, e_int, f_int = t.data[k]
        c.set_structure(r, elem, box)
        t.ctx.set_structure(r, elem, box)
        a, b = c.evaluate(1), t.ctx.evaluate(1)
        assert abs(a - b) <= 1e-12 * abs(b), (k, a, b)
