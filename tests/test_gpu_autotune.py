"""The launch-form tuning (hdnnp_ctx_autotune / hdnnp_ctx_set_tuning): SF
members per group chunk of the device tables (4, 8, 16), waves-per-SIMD
bounds of the specialised SF and force kernels and 16-atom tiles per wave
of the MFMA networks (1, 2, 4).

None of the knobs changes what is computed, only the order of some sums:
every setting is checked against the FP64 oracle with the bound of its
arithmetic (FP64: energy 1e-12 relative, forces 1e-9 eV/A; mixed: energy
1e-6 relative (FP16 networks 1e-5, BF16 networks 1e-4), force RMSE 1e-4
and largest error 1e-3 eV/A); the FP64 G
of every setting within 1e-12 (relative to the largest |G|) of the
oracle's, as the member sums are reduced in an order that depends on the
chunking.  The sweep returns its rows and marks the choice; a second
context of the same model adopts the cached choice without sweeping.
Cases: water (G2 + G3, groups of 8 members) and three elements with G2 +
G9 (groups of 10 and 12: three chunks of 4 per element pair); a batch of
structures of the second model (hdnnp_set_structures), whose settings and
sweep must give the default setting's FP64 results within 1e-12 (energies,
relative) and 1e-10 eV/A.
"""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import workloads as W
from tests import cases
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-tune]", *a, flush=True)
    sys.stdout.flush()


def _oracle(d, r, e, box):
    s = OStruct(OModel.from_dir(d), r, e, box)
    return s.evaluate(cells=True), s.forces(), s


# Energy bound (relative) per arithmetic; the mixed forms share the force
# bounds (RMSE 1e-4, largest error 1e-3 eV/A).
E_REL = {engine.PRECISION_MIXED: 1e-6, engine.PRECISION_MIXED16: 1e-5,
         engine.PRECISION_MIXED_BF16: 1e-4}


def _check(prec, en, f, oe, of):
    err = f - of
    if prec == engine.PRECISION_FP64:
        assert abs(en - oe) <= 1e-12 * abs(oe)
        assert np.abs(err).max() <= 1e-9
    else:
        rmse = float(np.sqrt(np.mean(err ** 2)))
        say("  E rel %.2e, F rmse %.2e max %.2e"
            % (abs(en - oe) / abs(oe), rmse, np.abs(err).max()))
        assert abs(en - oe) <= E_REL[prec] * abs(oe)
        assert rmse <= 1e-4 and np.abs(err).max() <= 1e-3


@pytest.mark.parametrize("case", ["water", "g9"])
def test_every_setting_matches_oracle(workdir, case):
    if case == "water":
        d, r, e, box = cases.water_case(workdir)
    else:
        d, r, e, box = cases.mixed_case(workdir, natoms=96, seed=3)
    oe, of, s = _oracle(d, r, e, box)
    g_ref = s.G()
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    for prec in (engine.PRECISION_FP64, engine.PRECISION_MIXED,
                 engine.PRECISION_MIXED16, engine.PRECISION_MIXED_BF16):
        ctx.set_precision(prec)
        for chunk, sfw, fw, tiles in ((16, -1, -1, 1), (8, 0, 0, 2),
                                      (4, 7, 4, 4), (4, 0, 2, 1),
                                      (8, 8, 6, 4)):
            say("%s prec %d chunk %d waves %d/%d tiles %d"
                % (case, prec, chunk, sfw, fw, tiles))
            ctx.set_tuning(chunk, sfw, fw, tiles)
            assert ctx.tuning() == (chunk, sfw, fw, tiles)
            en = ctx.evaluate(engine.FORCES_CONSISTENT)
            f = ctx.forces()
            _check(prec, en, f, oe, of)
            if prec == engine.PRECISION_FP64:
                G = ctx.G()
                assert G.shape == g_ref.shape
                assert np.abs(G - g_ref).max() <= 1e-12 * np.abs(g_ref).max()
    ctx.close()


def test_sweep_and_cache(workdir):
    d, r, e, box = cases.water_case(workdir)
    oe, of, _ = _oracle(d, r, e, box)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_precision(engine.PRECISION_MIXED)
    ctx.set_structure(r, e, box)
    say("sweep, water 192, mixed")
    rows = ctx.autotune(sweep=True, reps=2)
    say("%d rows, choice %s" % (len(rows), ctx.tuning()))
    assert len(rows) >= 4
    assert np.all(rows[:, 4:7] > 0.0)
    chosen = rows[:, 7].astype(int)
    assert np.any(chosen & 1) and np.any(chosen & 2)
    t = ctx.tuning()
    assert t[0] in (4, 8, 16) and t[3] in (1, 2, 4)
    # Profilers list the chosen launch form under its own kernel names.
    for which, w in (("sf", t[1]), ("force", t[2])):
        name = ctx.kernel_name(which)
        say("%s kernel: %s" % (which, name))
        assert name.startswith("hdnnp_plan_%s_f32_m" % which)
        assert name.endswith("_w%d" % w)
    en = ctx.evaluate(engine.FORCES_CONSISTENT)
    _check(engine.PRECISION_MIXED, en, ctx.forces(), oe, of)
    # A second context of the same model shape adopts the cached choice
    # without a structure and without sweeping.
    m2 = engine.Model.load_dir(d)
    ctx2 = engine.Context(m2)
    ctx2.set_precision(engine.PRECISION_MIXED)
    assert len(ctx2.autotune()) == 0
    assert ctx2.tuning() == t
    ctx2.set_structure(r, e, box)
    en2 = ctx2.evaluate(engine.FORCES_CONSISTENT)
    assert en2 == en
    assert np.array_equal(ctx2.forces(), ctx.forces())
    ctx2.close()
    ctx.close()


def test_batch_settings_and_sweep(workdir):
    d, _, _, _ = cases.mixed_case(workdir, natoms=96, seed=3)
    sts = [W.random_structure(nat, seed=300 + k)
           for k, nat in enumerate((16, 40, 96, 150, 23))]
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structures(sts)
    ctx.evaluate(engine.FORCES_REFERENCE)
    e0, f0 = ctx.structure_energies().copy(), ctx.forces().copy()
    g0 = ctx.G().copy()
    for t in ((4, 0, 2, 1), (8, 7, 4, 2), (16, 4, 0, 4)):
        say("batch fp64 setting %s" % (t,))
        ctx.set_tuning(*t)
        ctx.evaluate(engine.FORCES_REFERENCE)
        assert np.all(np.abs(ctx.structure_energies() - e0)
                      <= 1e-12 * np.abs(e0))
        assert np.abs(ctx.forces() - f0).max() <= 1e-10
        assert np.abs(ctx.G() - g0).max() <= 1e-12 * np.abs(g0).max()
    rows = ctx.autotune(engine.FORCES_REFERENCE, sweep=True, reps=2)
    say("batch fp64 sweep: %d rows, choice %s" % (len(rows), ctx.tuning()))
    assert len(rows) >= 4
    ctx.evaluate(engine.FORCES_REFERENCE)
    assert np.abs(ctx.forces() - f0).max() <= 1e-10
    ctx.set_precision(engine.PRECISION_MIXED)
    rows = ctx.autotune(engine.FORCES_REFERENCE, sweep=True, reps=2)
    say("batch mixed sweep: %d rows, choice %s" % (len(rows), ctx.tuning()))
    ctx.evaluate(engine.FORCES_REFERENCE)
    err = ctx.forces() - f0
    assert float(np.sqrt(np.mean(err ** 2))) <= 1e-4
    assert np.abs(err).max() <= 1e-3
    assert np.all(np.abs(ctx.structure_energies() - e0)
                  <= 1e-6 * np.abs(e0))
    ctx.close()


def test_tuning_rejects_bad_values(workdir):
    d, r, e, box = cases.water_case(workdir)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    for bad in ((0, -1, -1, 1), (17, -1, -1, 1), (8, 9, -1, 1),
                (8, -1, -2, 1), (8, -1, -1, 3)):
        with pytest.raises(RuntimeError):
            ctx.set_tuning(*bad)
    assert ctx.tuning() == (16, -1, -1, 1)
    # No structure: a sweep has nothing to time.
    with pytest.raises(RuntimeError, match="No structure"):
        ctx.autotune(sweep=True)
    ctx.close()

