"""Model-specialised kernels (compiled at run time with the member
constants folded in) against the generic kernels that read the member
tables, through the C ABI.

Both evaluate the same expressions in the same order, so the plain FP64
results agree to rounding: G and dE/dG within 1e-13 of their scale, energy
within 1e-13 relative, forces within 1e-12 eV/A.  On the mixed path (FP32
symmetry functions and force contraction) the two FP32 evaluations may
round differently; the bounds are those the mixed path meets against the
FP64 oracle: G and dE/dG within 1e-5 of their scale, energy within 1e-6
relative, forces within 1e-4 eV/A (observed on the Cu2S example: 1.3e-6,
2.2e-7 and 6e-6).  Every case first asserts that the specialised kernels really
compiled, so the comparison is never generic against generic.
"""
import os
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd.predict import Prediction
from tests import cases

pytestmark = pytest.mark.gpu

TOL = {engine.PRECISION_FP64: dict(g=1e-13, e=1e-13, f=1e-12),
       engine.PRECISION_MIXED: dict(g=1e-5, e=1e-6, f=1e-4)}


def say(*a):
    print("[gpu-spec]", *a, flush=True)
    sys.stdout.flush()


def _state(ctx, mode):
    e = ctx.evaluate(mode)
    return e, ctx.G(), ctx.dEdG(), ctx.forces()


def _compare(ctx, prec, mode=0):
    ctx.set_precision(prec)
    ctx.set_specialized(True)
    assert ctx.specialized_active(), "specialised kernels did not compile"
    e1, G1, D1, F1 = _state(ctx, mode)
    ctx.set_specialized(False)
    assert not ctx.specialized_active()
    e0, G0, D0, F0 = _state(ctx, mode)
    ctx.set_specialized(True)
    t = TOL[prec]
    gs = np.maximum(np.abs(G0).max(0), 1e-300)
    ds = max(np.abs(D0).max(), 1e-300)
    err = dict(g=float(np.max(np.abs(G1 - G0) / gs)),
               dEdG=float(np.max(np.abs(D1 - D0)) / ds),
               e=abs(e1 - e0) / max(abs(e0), 1e-300),
               f=float(np.abs(F1 - F0).max()))
    say("  precision", prec, "mode", mode,
        {k: "%.3g" % v for k, v in err.items()})
    assert err["g"] <= t["g"] and err["dEdG"] <= t["g"]
    assert err["e"] <= t["e"]
    assert err["f"] <= t["f"]


@pytest.mark.parametrize("prec", [engine.PRECISION_FP64,
                                  engine.PRECISION_MIXED])
def test_specialized_water192(workdir, prec):
    say("water192 specialised vs generic, precision", prec)
    d, r, e, box = cases.water_case(workdir)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    for mode in (0, 1):
        _compare(ctx, prec, mode)


@pytest.mark.parametrize("prec", [engine.PRECISION_FP64,
                                  engine.PRECISION_MIXED])
def test_specialized_g9_poly_softplus(workdir, prec):
    say("mixed 96 atoms (G2 + G9, poly2) specialised vs generic", prec)
    d, r, e, box = cases.mixed_case(workdir, natoms=96, seed=3)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    _compare(ctx, prec)


@pytest.mark.parametrize("name", ["H2O_RPBE-D3", "Cu2S_PBE"])
def test_specialized_examples(name):
    """The reference's example models: groups with more than eight members
    and several element pairs per angular group."""
    say("example", name, "specialised vs generic")
    p = Prediction(os.path.join(cases.EXAMPLES, name))
    p.setup()
    p.read_structure_from_file()
    p.predict()
    for prec in TOL:
        _compare(p.ctx, prec)


def test_specialized_non_integer_zeta(workdir):
    """Angular members with non-integer zeta, zeta = 1 and lambda = -1 mixed
    in one group (generic pow inside the specialised code)."""
    say("non-integer zeta group specialised vs generic")
    lines = ["H 3 H H 0.01 1.0 1.0 6.0", "H 3 H H 0.01 -1.0 2.5 6.0",
             "H 3 H H 0.01 1.0 3.7 6.0", "H 3 H H 0.01 -1.0 4.0 6.0",
             "H 9 H H 0.02 1.0 1.5 6.0", "H 9 H H 0.02 -1.0 8.0 6.0"]
    d = cases.single_sf_model(workdir, "gspec_zeta", "H", lines,
                              cutoff_type=1, act="t l", all_sf=True)
    m = engine.Model.load_dir(d, scaling=False)
    ctx = engine.Context(m)
    rng = np.random.default_rng(11)
    r = rng.uniform(0, 9.0, (60, 3))
    ctx.set_structure(r, np.zeros(60, dtype=np.int32), np.diag([9.0] * 3))
    for prec in TOL:
        _compare(ctx, prec)
