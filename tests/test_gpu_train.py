"""Training kernels of the radial-only surrogate (hdnnp_train_*).

- The loss gradient against central finite differences (h = 1e-4) of the
  same loss computed by the oracle (energies and forces of the surrogate
  model with perturbed weight files): error at most 1e-6 of the larger of
  |FD| and 1e-3 of the largest gradient component, on weights of every
  layer, biases included, of both elements.  The loss itself agrees to
  1e-10.  The oracle knows nothing of the
  gradient, so a wrong term, sign or index in the forward-over-reverse
  derivation fails here.
- One Adam step against the published update rule (Kingma & Ba 2015,
  Algorithm 1) evaluated in numpy on the same gradient: 1e-14 relative.
- Training recovers a teacher of the same architecture (force loss down
  100x), and distilling the full model lowers its losses.
"""
import os
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import surrogate as S
from n2p2_amd import workloads as W
from tests import cases
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-train]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sur(tmp_path_factory):
    base = tmp_path_factory.mktemp("train")
    d, r, e, box = cases.water_case(base)
    sd = os.path.join(str(base), "surrogate")
    S.write_model_gpu_scaled(sd, S.surrogate_settings(), W.WATER_ELEMENTS,
                             (1, 8), S.SURROGATE_HIDDEN, r, e, box, seed=21)
    return d, sd, r, e, box


def _oracle_loss(sd, w, r, e, box, fref, eref, w_e, w_f, nsf):
    S.write_weights(sd, w, nsf, (1, 8))
    om = OModel.from_dir(sd)
    s = OStruct(om, r, e, box)
    en = s.evaluate(cells=True, force_mode=1)
    n = len(e)
    df = s.forces() - fref
    return w_e * ((en - eref) / n) ** 2 + w_f * np.sum(df ** 2) / (3 * n)


def test_gradient_vs_oracle_finite_differences(sur, tmp_path):
    d, sd0, r, e, box = sur
    sd = str(tmp_path / "fd")
    os.makedirs(sd)
    for f in ("input.nn", "scaling.data"):
        with open(os.path.join(sd0, f)) as a, open(os.path.join(sd, f), "w") as b:
            b.write(a.read())
    m = engine.Model.load_dir(sd0)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    rng = np.random.default_rng(3)
    w0 = ctx.get_weights()
    # nonzero biases so that their gradients are exercised too
    w0 = w0 + rng.normal(scale=0.05, size=w0.size)
    ctx.set_weights(w0)
    fref = rng.normal(scale=0.3, size=(len(e), 3))
    eref = -5.0
    w_e, w_f = 0.7, 1.3
    g, loss = ctx.train_gradient(fref, eref, w_e, w_f)
    nsf = [m.num_sf(k) for k in range(m.num_elements)]
    lo = _oracle_loss(sd, w0, r, e, box, fref, eref, w_e, w_f, nsf)
    say("loss gpu %.12g + %.12g, oracle %.12g" % (loss[0], loss[1], lo))
    assert abs(loss[0] + loss[1] - lo) <= 1e-10 * abs(lo)
    # weights of every layer and bias of both elements
    nw0 = sum(a * b + b for a, b in zip(S.layer_sizes(nsf[0])[:-1],
                                        S.layer_sizes(nsf[0])[1:]))
    picks = [0, 5, 8 * 16 + 3, 8 * 16 + 16 + 17, 8 * 16 + 16 + 256 + 2,
             nw0 - 16, nw0 - 1, nw0 + 1, nw0 + 100, len(w0) - 1]
    h = 1e-4
    gs = np.abs(g).max()
    for k in picks:
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        fd = (_oracle_loss(sd, wp, r, e, box, fref, eref, w_e, w_f, nsf)
              - _oracle_loss(sd, wm, r, e, box, fref, eref, w_e, w_f, nsf)) \
            / (2 * h)
        rel = abs(g[k] - fd) / max(abs(fd), 1e-3 * gs)
        say("  weight %d: gpu %.10g fd %.10g rel %.2g" % (k, g[k], fd, rel))
        assert rel <= 1e-6


@pytest.mark.parametrize("which", ["water_g2_g3_tanh", "mixed_g2_g9_softplus",
                                   "weighted_g12_g13_tanh",
                                   "compact_g20_g25_tanh"])
def test_gradient_full_models_vs_oracle_fd(which, tmp_path):
    """The same check on the full models with angular SFs: water (G2 + G3,
    tanh), the three-element set (G2 + G9, softplus), the set with the
    element-weighted types 12 and 13 and the set with the compact types
    20-25."""
    if which.startswith("water"):
        d, r, e, box = cases.water_case(tmp_path / "w")
        z, hidden = (1, 8), (25, 25)
    elif which.startswith("weighted"):
        d, r, e, box = cases.weighted_case(tmp_path / "g", natoms=40, seed=3)
        z, hidden = (5, 6, 7), (20, 20)
    elif which.startswith("compact"):
        d, r, e, box = cases.compact_case(tmp_path / "c", natoms=40, seed=3)
        z, hidden = (5, 6, 7), (20, 20)
    else:
        d, r, e, box = cases.mixed_case(tmp_path / "m", natoms=40, seed=3)
        z, hidden = (5, 6, 7), (30, 30)
    sd = str(tmp_path / "fd")
    os.makedirs(sd)
    for f in ("input.nn", "scaling.data"):
        with open(os.path.join(d, f)) as a, open(os.path.join(sd, f), "w") as b:
            b.write(a.read())
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    rng = np.random.default_rng(8)
    w0 = ctx.get_weights() + rng.normal(scale=0.02, size=ctx.num_weights())
    ctx.set_weights(w0)
    fref = rng.normal(scale=0.3, size=(len(e), 3))
    eref, w_e, w_f = -3.0, 0.5, 1.5
    g, loss = ctx.train_gradient(fref, eref, w_e, w_f)
    nsf = [m.num_sf(k) for k in range(m.num_elements)]

    def oloss(w):
        S.write_weights(sd, w, nsf, z, hidden)
        s = OStruct(OModel.from_dir(sd), r, e, box)
        en = s.evaluate(cells=True, force_mode=1)
        df = s.forces() - fref
        n = len(e)
        return w_e * ((en - eref) / n) ** 2 + w_f * np.sum(df ** 2) / (3 * n)

    lo = oloss(w0)
    say("%s: loss gpu %.12g oracle %.12g" % (which, loss[0] + loss[1], lo))
    assert abs(loss[0] + loss[1] - lo) <= 1e-10 * abs(lo)
    h = 1e-4
    gs = np.abs(g).max()
    picks = list(rng.choice(len(w0), 8, replace=False)) + [0, len(w0) - 1]
    for k in picks:
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        fd = (oloss(wp) - oloss(wm)) / (2 * h)
        rel = abs(g[k] - fd) / max(abs(fd), 1e-3 * gs)
        say("  %s weight %d: gpu %.10g fd %.10g rel %.2g" % (which, k, g[k],
                                                             fd, rel))
        assert rel <= 1e-6


def test_adam_step_matches_update_rule(sur):
    d, sd, r, e, box = sur
    m = engine.Model.load_dir(sd)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    rng = np.random.default_rng(4)
    fref = rng.normal(scale=0.3, size=(len(e), 3))
    w = ctx.get_weights()
    mm = np.zeros_like(w)
    vv = np.zeros_like(w)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    ctx.train_reset()
    for t in (1, 2, 3):
        g, _ = ctx.train_gradient(fref, -3.0)
        ctx.train_adam(lr, b1, b2, eps)
        mm = b1 * mm + (1 - b1) * g
        vv = b2 * vv + (1 - b2) * g * g
        w = w - lr * (mm / (1 - b1 ** t)) / (np.sqrt(vv / (1 - b2 ** t)) + eps)
        wg = ctx.get_weights()
        err = np.max(np.abs(wg - w) / np.maximum(np.abs(w), 1e-3))
        say("adam step %d: max rel diff %.2g" % (t, err))
        assert err <= 1e-14


def test_training_recovers_a_radial_teacher(sur, tmp_path):
    """Teacher and student share the surrogate architecture and
    descriptors, with different random weights: training on the teacher's
    MD frames must drive the force loss down by two orders of magnitude,
    which only a correct gradient and update achieve."""
    d, sd, r, e, box = sur
    td = str(tmp_path / "teacher")
    teacher = S.write_model_gpu_scaled(td, S.surrogate_settings(),
                                       W.WATER_ELEMENTS, (1, 8),
                                       S.SURROGATE_HIDDEN, r, e, box, seed=77)
    masses = np.array(W.WATER_MASSES)
    R, E, F = S.generate_frames(teacher, r, e, box, masses, nframes=16,
                                stride=8)
    # the student uses the teacher's scaling, so the targets are reachable
    sd2 = str(tmp_path / "student")
    os.makedirs(sd2)
    for f in ("input.nn", "scaling.data"):
        with open(os.path.join(td, f)) as a, open(os.path.join(sd2, f), "w") as b:
            b.write(a.read())
    W.write_model(sd2, open(os.path.join(td, "input.nn")).read(),
                  W.WATER_ELEMENTS, (1, 8), S.SURROGATE_HIDDEN, seed=5)
    ctx = engine.Context(engine.Model.load_dir(sd2))
    hist = S.train(ctx, e, box, R, E, F, epochs=40, lr=5e-3, log=say)
    say("teacher-student force loss %.4g -> %.4g" % (hist[0][1], hist[-1][1]))
    assert hist[-1][1] < 0.01 * hist[0][1]


def test_distillation_of_the_full_model(sur):
    """The radial-only surrogate on frames of the full (angular) model's own
    MD: energy loss falls by an order of magnitude, force loss by a quarter
    (a radial descriptor cannot represent the angular part)."""
    d, sd, r, e, box = sur
    full = engine.Model.load_dir(d)
    masses = np.array(W.WATER_MASSES)
    R, E, F = S.generate_frames(full, r, e, box, masses, nframes=24,
                                stride=8)
    m = engine.Model.load_dir(sd)
    ctx = engine.Context(m)
    hist = S.train(ctx, e, box, R, E, F, epochs=20, lr=2e-3, log=say)
    say("force loss %.4g -> %.4g" % (hist[0][1], hist[-1][1]))
    assert hist[-1][0] < 0.1 * hist[0][0]
    assert hist[-1][1] < 0.75 * hist[0][1]
