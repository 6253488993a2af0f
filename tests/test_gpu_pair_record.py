"""The pair record of the mixed path (hdnnp_ctx_set_pair_record): the
symmetry-function kernel records every angular group's passing neighbour
pairs and the force kernel replays them instead of enumerating and testing
the pairs again.

The replay hands every lane the triplets the walk would have handed it, in
the same order, so the results are the same bit for bit with and without
the record: energy, forces, G and dE/dG compared with np.array_equal.
Against the FP64 oracle the mixed-path bounds hold (energy 1e-6 relative,
force RMSE 1e-4, largest error 1e-3 eV/A).  Cases: water (G2 + G3, cosine
cutoff, tanh), three elements with G2 + G9 (polynomial cutoff, softplus),
the weighted types, neighbour lists longer than 256 (staged indices do not
fit a byte: no record, both kernels walk) and more passing pairs than the
record holds (the groups past it walk again).  Also: MD steps with the
record, captured in graphs, equal eager ones bit for bit.
"""
import os
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import workloads as W
from tests import cases
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-record]", *a, flush=True)
    sys.stdout.flush()


def _eval(m, r, e, box, record):
    ctx = engine.Context(m)
    ctx.set_precision(engine.PRECISION_MIXED)
    ctx.set_pair_record(record)
    ctx.set_structure(r, e, box)
    en = ctx.evaluate(engine.FORCES_REFERENCE)
    active = ctx.pair_record_active()
    out = dict(e=en, f=ctx.forces(), G=ctx.G(), D=ctx.dEdG(),
               ea=ctx.atomic_energies(), active=active)
    ctx.close()
    return out


def _compare(d, r, e, box, expect_record=True):
    m = engine.Model.load_dir(d)
    a = _eval(m, r, e, box, True)
    b = _eval(m, r, e, box, False)
    assert a["active"] == expect_record and not b["active"]
    assert a["e"] == b["e"]
    for k in ("f", "G", "D", "ea"):
        assert np.array_equal(a[k], b[k]), k
    s = OStruct(OModel.from_dir(d), r, e, box)
    oe = s.evaluate(cells=True)
    of = s.forces()
    err = a["f"] - of
    rmse = float(np.sqrt(np.mean(err ** 2)))
    say("  record on = off bitwise; vs oracle: E %.2e, F rmse %.2e max %.2e"
        % (abs(a["e"] - oe) / abs(oe), rmse, np.abs(err).max()))
    assert abs(a["e"] - oe) <= 1e-6 * abs(oe)
    assert rmse <= 1e-4 and np.abs(err).max() <= 1e-3


def test_record_water(workdir):
    say("water 192")
    d, r, e, box = cases.water_case(workdir)
    _compare(d, r, e, box)


def test_record_g9_three_elements(workdir):
    say("B/C/N 96 atoms, G2 + G9")
    d, r, e, box = cases.mixed_case(workdir, natoms=96, seed=3)
    _compare(d, r, e, box)


def test_record_weighted(workdir):
    say("weighted types 12 and 13")
    d, r, e, box = cases.weighted_case(workdir, natoms=96, seed=5)
    _compare(d, r, e, box)


def test_record_with_compact_groups(workdir):
    """The compact families run in their own pass beside the recorded
    exponential groups."""
    say("compact types 20-25")
    d, r, e, box = cases.compact_case(workdir, natoms=64, seed=7)
    _compare(d, r, e, box)


def _model(tmp, name, settings, elements, z, hidden, r, e, box):
    d = os.path.join(str(tmp), name)
    if not os.path.exists(os.path.join(d, "scaling.data")):
        W.write_model(d, settings, elements, z, hidden, seed=17)
        cases.scale_from_oracle(d, r, e, box)
    return d


def test_record_long_lists(workdir):
    """rc 9 A on water: about 300 neighbours, so staged indices do not
    fit a byte: no record, both kernels walk."""
    r, e, box = W.water(216, 18.63, seed=31)
    d = _model(workdir, "water_rc9", W.water_settings(rc=9.0, hidden=(20, 20)),
               W.WATER_ELEMENTS, (1, 8), (20, 20), r, e, box)
    say("water 648, rc 9")
    _compare(d, r, e, box, expect_record=False)


def test_record_overflow(workdir):
    """G9 with rc 8 A at 0.08 atoms/A^3: about 170 neighbours and more
    passing pairs (all of them, wide groups) than the 4096 recorded, so the
    later groups walk again."""
    r, e, box = W.random_structure(400, seed=33)
    d = _model(workdir, "g9_rc8", W.mixed_settings(rc=8.0),
               W.MIXED_ELEMENTS, (5, 6, 7), (30, 30), r, e, box)
    say("B/C/N 400 atoms, G9, rc 8")
    _compare(d, r, e, box)


def test_record_md_graph_equals_eager_and_repeats(workdir):
    d, r, e, box = cases.water_case(workdir)
    m = engine.Model.load_dir(d)
    masses = np.array(W.WATER_MASSES)
    v = W.maxwell_boltzmann(masses[e], 300.0, seed=5)
    out = []
    for graph in (True, False):
        ctx = engine.Context(m)
        ctx.set_precision(engine.PRECISION_MIXED)
        ctx.set_structure(r, e, box)
        ctx.md_init(masses, 0.5, skin=1.0, v=v, graph=graph)
        assert ctx.pair_record_active()
        ep, ek = ctx.md_run(60)
        x, u = ctx.md_state()
        out.append((ep, ek, x, u, ctx.md_rebuilds(), ctx.md_partner_misses()))
        ctx.close()
    say("MD with the record: rebuilds %d" % out[0][4])
    assert out[0][4] >= 2 and out[0][5] == 0
    for p, q in zip(out[0][:4], out[1][:4]):
        assert np.array_equal(p, q)
    a = _eval(m, r, e, box, True)
    b = _eval(m, r, e, box, True)
    assert a["e"] == b["e"] and np.array_equal(a["f"], b["f"])
