"""Mixed-precision paths against the FP64 oracle.

PRECISION_MIXED: FP32 symmetry functions and force contraction, FP32 MFMA
network layers; FP64 scaling, biases, activations, energy total and force
accumulation.  PRECISION_MIXED16 / PRECISION_MIXED_BF16: the same with FP16
/ BF16 MFMA network layers, every FP32 operand split into a rounded hi
part and a rounded lo remainder (three MFMAs per step, FP32 accumulation),
so the operands carry 22 / 16 significant bits.

The bound the paths are built to is a force RMSE of at most 1 meV/A against
FP64.  For MIXED the tests assert ten times tighter RMSE (1e-4 eV/A), the
1 meV/A bound on the largest component error, energy within 1e-6 relative
and G within 1e-5 of each SF's scale.  MIXED16 (observed: energy 2e-8 to
2e-7 relative, force RMSE 3e-7) is held to the same force bounds and 1e-5
on the energy; MIXED_BF16 (observed: energy up to 6.3e-6 relative, force
RMSE 4.8e-6, largest error 2e-5 eV/A) to the same force bounds and 1e-4
on the energy.  A wrong operand layout or a dropped lo term exceeds these
by orders of magnitude.  Every test prints the errors it saw.
"""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import workloads as W
from tests import cases
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu

BOUNDS = {engine.PRECISION_MIXED: dict(e_rel=1e-6, f_rmse=1e-4, f_max=1e-3,
                                       g_rel=1e-5),
          engine.PRECISION_MIXED16: dict(e_rel=1e-5, f_rmse=1e-4,
                                         f_max=1e-3, g_rel=1e-5),
          engine.PRECISION_MIXED_BF16: dict(e_rel=1e-4, f_rmse=1e-4,
                                            f_max=1e-3, g_rel=1e-5)}
PRECISIONS = [engine.PRECISION_MIXED, engine.PRECISION_MIXED16,
              engine.PRECISION_MIXED_BF16]


def say(*a):
    print("[gpu-mixed]", *a, flush=True)
    sys.stdout.flush()


def _mixed_ctx(d, scaling=True, prec=engine.PRECISION_MIXED):
    m = engine.Model.load_dir(d, scaling=scaling)
    ctx = engine.Context(m)
    ctx.set_precision(prec)
    return m, ctx


def _errors(ctx, s, e, oe):
    f, of = ctx.forces(), s.forces()
    df = f - of
    G, oG = ctx.G(), s.G()
    scale = np.maximum(np.abs(oG).max(0), 1e-300)
    out = {"e_rel": abs(e - oe) / abs(oe),
           "f_rmse": float(np.sqrt(np.mean(df ** 2))),
           "f_max": float(np.abs(df).max()),
           "f_scale": float(np.sqrt(np.mean(of ** 2))),
           "g_rel": float(np.max(np.abs(G - oG) / scale))}
    say("errors vs FP64 oracle:", {k: "%.3g" % v for k, v in out.items()})
    return out


def _check(err, prec):
    b = BOUNDS[prec]
    for k in ("e_rel", "f_rmse", "f_max", "g_rel"):
        assert err[k] <= b[k], (k, err[k], b[k])


@pytest.mark.parametrize("prec", PRECISIONS)
def test_mixed_water192(workdir, prec):
    d, r, e, box = cases.water_case(workdir)
    m, ctx = _mixed_ctx(d, prec=prec)
    ctx.set_structure(r, e, box)
    s = OStruct(OModel.from_dir(d), r, e, box)
    for mode in (0, 1):
        say("water192, precision", prec, "force mode", mode)
        en = ctx.evaluate(mode)
        oe = s.evaluate(cells=True, force_mode=mode)
        _check(_errors(ctx, s, en, oe), prec)


@pytest.mark.parametrize("prec", PRECISIONS)
def test_mixed_g9_poly_softplus(workdir, prec):
    d, r, e, box = cases.mixed_case(workdir, natoms=96, seed=3)
    say("mixed 96 atoms (G2 + G9, poly2, softplus), precision", prec)
    m, ctx = _mixed_ctx(d, prec=prec)
    ctx.set_structure(r, e, box)
    s = OStruct(OModel.from_dir(d), r, e, box)
    en = ctx.evaluate(0)
    oe = s.evaluate(cells=True)
    _check(_errors(ctx, s, en, oe), prec)


@pytest.mark.parametrize("ct", range(9))
def test_mixed_cutoff_types(workdir, ct):
    """Two-atom G2 model per cutoff type: FP32 cutoff value and derivative
    against the reference's own test values (test_CutoffFunction.cpp)."""
    say("mixed cutoff type", ct)
    g = cases.load_json("cutoff_functions.json")
    d = cases.single_sf_model(workdir, "mcut%d" % ct, "H", ["H 2 H 0.0 0.0 3.0"],
                              cutoff_type=ct, cutoff_alpha=g["alpha"])
    m, ctx = _mixed_ctx(d, scaling=False)
    ctx.set_structure(np.array([[0.0, 0.0, 0.0], [g["r"], 0.0, 0.0]]), [0, 0])
    e = ctx.evaluate(0)
    f = ctx.forces()
    say("  G %.9g vs %.9g, F %.9g vs %.9g" % (ctx.G()[0, 0], g["f"][ct],
                                             f[0, 0], 2 * g["df"][ct]))
    assert abs(ctx.G()[0, 0] - g["f"][ct]) <= 1e-6
    assert abs(e - 2 * g["f"][ct]) <= 2e-6
    assert abs(f[0, 0] - 2 * g["df"][ct]) <= 2e-6 * max(1.0, abs(g["df"][ct]))


def test_mixed_water24576(workdir):
    say("water 24,576 atoms, both mixed precisions vs FP64 oracle")
    d, _, _, _ = cases.water_case(workdir)
    r, e, box = W.water(8192, 62.1, 99)
    s = OStruct(OModel.from_dir(d), r, e, box)
    oe = s.evaluate(cells=True)
    for prec in PRECISIONS:
        say("  precision", prec)
        m, ctx = _mixed_ctx(d, prec=prec)
        ctx.set_structure(r, e, box)
        en = ctx.evaluate(0)
        _check(_errors(ctx, s, en, oe), prec)


def test_mixed_deterministic_and_switch(workdir):
    """Repeated mixed evaluations are bitwise equal; switching back to FP64
    recomputes everything and matches the oracle at the FP64 tolerance."""
    say("mixed determinism and precision switch")
    d, r, e, box = cases.water_case(workdir)
    m, ctx = _mixed_ctx(d)
    ctx.set_structure(r, e, box)
    e1 = ctx.evaluate(0)
    f1 = ctx.forces()
    e2 = ctx.evaluate(0)
    f2 = ctx.forces()
    assert e1 == e2 and np.array_equal(f1, f2)
    ctx.set_precision(engine.PRECISION_FP64)
    e3 = ctx.evaluate(0)
    s = OStruct(OModel.from_dir(d), r, e, box)
    oe = s.evaluate(cells=True)
    assert abs(e3 - oe) <= 1e-10 * abs(oe)
    assert np.abs(ctx.forces() - s.forces()).max() < 1e-9
