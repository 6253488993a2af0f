"""Kalman filter on the weights (hdnnp_kalman_*, KalmanFilter.cpp).

- With a fixed learning rate (etatau = 0) and no process noise (q = 0), a
  sequence of scalar updates is recursive least squares with prior
  precision epsilon I and measurement precision eta, so after linear
  observations y_k = h_k . w_true the weights equal the closed form
      w = w0 + (eps I + eta H^T H)^-1 eta H^T (y - H w0)
  (a numpy solve): relative error 1e-9.  The filter's P update, gain and
  weight update all enter; a wrong sign, factor or transposition fails.
- The trainer's Kalman updater (energy updates and a fraction of force
  updates, the force Jacobian from the loss gradient) lowers the training
  RMSE on the reference's example data set.
"""
import os
import shutil
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import surrogate as S
from n2p2_amd import workloads as W
from n2p2_amd.train import Training
from tests import cases

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "nnp_scaling_H2O")


def say(*a):
    print("[gpu-kalman]", *a, flush=True)
    sys.stdout.flush()


def test_kalman_equals_recursive_least_squares(tmp_path):
    d, r, e, box = cases.water_case(tmp_path / "w")
    sd = str(tmp_path / "sur")
    S.write_model_gpu_scaled(sd, S.surrogate_settings(), W.WATER_ELEMENTS,
                             (1, 8), S.SURROGATE_HIDDEN, r, e, box, seed=3)
    ctx = engine.Context(engine.Model.load_dir(sd))
    nw = ctx.num_weights()
    rng = np.random.default_rng(12)
    w0 = ctx.get_weights()
    w_true = w0 + rng.normal(scale=0.1, size=nw)
    eps, eta, M = 0.5, 3.0, 200
    ctx.kalman_init(epsilon=eps, q0=0.0, qtau=0.0, qmin=0.0, eta0=eta,
                    etatau=0.0, etamax=eta)
    H = rng.normal(size=(M, nw))
    y = H @ w_true
    for k in range(M):
        w = ctx.get_weights()
        ctx.kalman_update(H[k], y[k] - H[k] @ w)
    w_kf = ctx.get_weights()
    A = eps * np.eye(nw) + eta * H.T @ H
    w_ls = w0 + np.linalg.solve(A, eta * H.T @ (y - H @ w0))
    err = np.max(np.abs(w_kf - w_ls)) / np.max(np.abs(w_ls - w0))
    say("nw %d, %d updates: max |w_kf - w_ls| / |w_ls - w0| = %.2e"
        % (nw, M, err))
    assert err <= 1e-9


def test_trainer_kalman_lowers_rmse(tmp_path):
    for f in ("input.nn", "input.data", "scaling.data"):
        shutil.copy(os.path.join(GOLD, f), tmp_path)
    t = Training(str(tmp_path))
    t.setup()
    assert t.updater == 1
    t.read_data()
    curve = t.run(epochs=2, log=say)
    assert curve[-1][1] < curve[0][1] and curve[-1][3] < curve[0][3]
