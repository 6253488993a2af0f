"""GPU parity: the HIP engine (through the C ABI) against the oracle and the
reference's goldens.

Tolerances (plain FP64 path): neighbour lists bit-exact (index, image, d,
dr); energy 1e-10 relative; forces 1e-9 eV/A absolute (physical units);
G and dE/dG 1e-11 relative to their scale (summation order differs from the
reference's sequential loops).
"""
import os
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import io as nio
from n2p2_amd import workloads as W
from n2p2_amd.predict import Prediction
from tests import cases
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu

E_REL = 1e-10
F_ABS = 1e-9
EPS = np.finfo(float).eps


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _engine(d, scaling=True):
    m = engine.Model.load_dir(d, scaling=scaling)
    return m, engine.Context(m)


def _neighbors_equal(ctx, s, r, box):
    off, idx, img, d, dr = ctx.neighbors()
    ooff, oidx, od, odr = s.neighbors()
    assert np.array_equal(off, ooff)
    assert np.array_equal(idx, oidx)
    assert np.array_equal(d, od), np.abs(d - od).max()
    assert np.array_equal(dr, odr)
    if box is not None:
        # image vectors reproduce dr exactly by the reference formula
        for q in range(0, len(idx), max(1, len(idx) // 500)):
            i = np.searchsorted(off, q, side="right") - 1
            j = idx[q]
            b = img[q]
            v = r[i] - r[j]
            v = v + b[0] * box[0]
            v = v + b[1] * box[1]
            v = v + b[2] * box[2]
            assert np.array_equal(v, dr[q])


def _small_cases(workdir):
    rng = np.random.default_rng(5)
    out = []
    r, e, box = W.water(64, 12.42, 1234)
    out.append(("water192", r, e, box, 6.35))
    box = np.diag([4.0, 4.5, 5.0])
    out.append(("small_box", rng.uniform(0, 4.0, (7, 3)), rng.integers(0, 2, 7), box, 6.35))
    box = np.array([[9.0, 0.0, 0.0], [2.0, 8.5, 0.0], [1.0, -1.5, 10.0]])
    out.append(("triclinic", nio.remap(rng.uniform(0, 9.0, (40, 3)), box),
                rng.integers(0, 2, 40), box, 4.0))
    out.append(("open", rng.uniform(0, 12.0, (50, 3)), rng.integers(0, 2, 50), None, 6.35))
    return out


def test_neighbor_list_bit_exact(workdir):
    d, _, _, _ = cases.water_case(workdir)
    m, ctx = _engine(d)
    om = OModel.from_dir(d)
    for name, r, e, box, rc in _small_cases(workdir):
        say("neighbor list", name)
        ctx.set_structure(r, e, box)
        ctx.calculate_neighbor_list(rc)
        s = OStruct(om, r, e, box)
        s.neighbor_list(rc)
        _neighbors_equal(ctx, s, r, box)
        if box is not None:
            assert np.array_equal(ctx.pbc(), s.pbc())


@pytest.mark.parametrize("name", cases.EXAMPLE_NAMES)
def test_example_prediction_goldens(name):
    say("example prediction", name)
    g = cases.load_json("example_prediction.json")["examples"][name]
    p = Prediction(os.path.join(cases.EXAMPLES, name))
    p.setup()
    p.read_structure_from_file()
    e, f = p.predict()
    f0 = np.array(g["forces"])
    assert abs(e - g["energy"]) <= E_REL * abs(g["energy"])
    assert np.abs(f - f0).max() <= F_ABS
    # Neighbour list of the normalised structure is bit-identical too.
    om = OModel.from_dir(os.path.join(cases.EXAMPLES, name))
    r, box = nio.to_normalized(p.r_phys, p.box_phys, p.conv_length)
    s = OStruct(om, r, p.elem, box)
    s.neighbor_list(om.max_cutoff)
    _neighbors_equal(p.ctx, s, r, box)


def _compare_full(ctx, s, force_mode=0, grel=1e-11):
    G, oG = ctx.G(), s.G()
    scale = np.maximum(np.abs(oG).max(0), 1e-300)
    assert np.max(np.abs(G - oG) / scale) < grel
    dEdG, odEdG = ctx.dEdG(), s.dEdG()
    scale = np.maximum(np.abs(odEdG).max(), 1e-300)
    assert np.max(np.abs(dEdG - odEdG)) / scale < grel
    assert np.allclose(ctx.atomic_energies(), s.atomic_energies(), rtol=1e-11,
                       atol=1e-13)
    f, of = ctx.forces(), s.forces()
    assert np.abs(f - of).max() < F_ABS


def _run_both(ctx, s, force_mode=0):
    e = ctx.evaluate(force_mode)
    oe = s.evaluate(cells=True, force_mode=force_mode)
    assert abs(e - oe) <= E_REL * abs(oe)
    return e, oe


def _pair_force_check(ctx, s, r, e):
    sf, pf = ctx.pair_forces()
    osf, opf = s.pair_forces()
    off, idx, _, _ = s.neighbors()
    # 1e-10 absolute per slot (forces are checked at 1e-9).  The cosine
    # cutoff's derivative is sin(pi x) here and sqrt(1 - cos^2) in the
    # reference; for pairs within ~1e-6 A of the cutoff the reference form
    # loses digits to cancellation (observed: 1.7e-11 on 2 of 2.6M slots).
    bad = np.nonzero(np.abs(pf - opf).max(1) > 1e-10)[0]
    if len(bad):
        q = bad[0]
        i = np.searchsorted(off, q, side="right") - 1
        ne = [int(np.sum(e[idx[off[i]:off[i + 1]]] == k)) for k in range(e.max() + 1)]
        msg = ("%d bad slots of %d; first: atom %d (element %d, %d neighbours, "
               "per element %s), slot %d -> atom %d (element %d): gpu %s oracle %s"
               % (len(bad), len(pf), i, e[i], off[i + 1] - off[i], ne, q - off[i],
                  idx[q], e[idx[q]], pf[q], opf[q]))
        raise AssertionError(msg)
    assert np.abs(sf - osf).max() <= 1e-10


def test_pair_forces_per_slot(workdir):
    """Atom::calculatePairForceShort of every slot and the self forces."""
    cases_ = []
    d = os.path.join(cases.EXAMPLES, "Cu2S_PBE")
    p = Prediction(d)
    p.setup()
    p.read_structure_from_file()
    r, box = nio.to_normalized(p.r_phys, p.box_phys, p.conv_length)
    cases_.append(("Cu2S", d, r, p.elem, box))
    dw, _, _, _ = cases.water_case(workdir)
    r2, e2, b2 = W.water(8192, 62.1, 99)
    cases_.append(("water24576", dw, r2, e2, b2))
    for name, dd, r, e, box in cases_:
        say("pair forces", name)
        m, ctx = _engine(dd)
        ctx.set_structure(r, e, box)
        ctx.evaluate(0)
        s = OStruct(OModel.from_dir(dd), r, e, box)
        s.evaluate(cells=True)
        _pair_force_check(ctx, s, r, np.asarray(e))


def test_water192_parity(workdir):
    d, r, e, box = cases.water_case(workdir)
    m, ctx = _engine(d)
    ctx.set_structure(r, e, box)
    s = OStruct(OModel.from_dir(d), r, e, box)
    for mode in (0, 1):
        say("water192 parity, force mode", mode)
        _run_both(ctx, s, mode)
        _compare_full(ctx, s, mode)


def test_mixed_g9_poly_softplus_parity(workdir):
    d, r, e, box = cases.mixed_case(workdir, natoms=96, seed=3)
    say("mixed 96 atoms (G2 + G9, poly2, softplus)")
    m, ctx = _engine(d)
    ctx.set_structure(r, e, box)
    s = OStruct(OModel.from_dir(d), r, e, box)
    _run_both(ctx, s, 0)
    _compare_full(ctx, s)


@pytest.mark.parametrize("ct", range(9))
def test_cutoff_goldens_gpu(workdir, ct):
    say("cutoff type", ct)
    g = cases.load_json("cutoff_functions.json")
    d = cases.single_sf_model(workdir, "gcut%d" % ct, "H", ["H 2 H 0.0 0.0 3.0"],
                              cutoff_type=ct, cutoff_alpha=g["alpha"])
    m, ctx = _engine(d, scaling=False)
    ctx.set_structure(np.array([[0.0, 0.0, 0.0], [g["r"], 0.0, 0.0]]), [0, 0])
    e = ctx.evaluate(0)
    # test_CutoffFunction.cpp tests each value at 100 eps absolute.
    eps = np.finfo(float).eps
    assert abs(ctx.G()[0, 0] - g["f"][ct]) <= 100 * eps
    assert abs(e - 2 * g["f"][ct]) <= 200 * eps
    f = ctx.forces()
    assert abs(f[0, 0] - 2 * g["df"][ct]) <= 200 * eps


def test_symfunc_fixture_values_gpu(workdir):
    fx = cases.load_json("symfunc_fixtures.json")
    for key, st in fx["sets"].items():
        syms = st["elements"].split()
        for k, sf in enumerate(st["sf"]):
            say("fixture", key, sf["line"])
            centre = sf["line"].split()[0]
            lines = [sf["line"]] + ["%s 2 %s 0.001 0.0 10.0" % (s, centre)
                                    for s in syms if s != centre]
            d = cases.single_sf_model(workdir, "gfx_%s_%d" % (key, k),
                                      st["elements"], lines, cutoff_type=2)
            m, ctx = _engine(d, scaling=False)
            ctx.set_structure(np.array(st["r"], dtype=float), st["elem"])
            ctx.evaluate(0)
            # The reference test's bound, 1000 eps absolute (the weighted
            # values reach 78 with uranium neighbours).
            assert abs(ctx.G()[0, 0] - sf["value"]) <= 1000 * EPS
            s = OStruct(OModel.from_dir(d, scaling=False), st["r"], st["elem"])
            s.evaluate()
            fo = s.forces()
            assert np.abs(ctx.forces() - fo).max() < 1e-12 * max(
                1.0, np.abs(fo).max())


def test_self_image_modes_gpu(workdir):
    say("self-image forces (cell smaller than cutoff)")
    d = cases.single_sf_model(
        workdir, "gselfimg", "H",
        ["H 2 H 0.01 0.0 6.0", "H 3 H H 0.01 1.0 1.0 6.0",
         "H 9 H H 0.01 -1.0 2.0 6.0"], cutoff_type=1, act="t l", all_sf=True)
    m, ctx = _engine(d, scaling=False)
    r = np.array([[0.3, 0.4, 0.2], [1.9, 1.1, 0.7], [0.8, 2.6, 1.9]])
    box = np.diag([5.0, 5.0, 5.0])
    ctx.set_structure(r, [0, 0, 0], box)
    s = OStruct(OModel.from_dir(d, scaling=False), r, [0, 0, 0], box)
    for mode in (0, 1):
        ctx.evaluate(mode)
        s.evaluate(force_mode=mode)
        assert np.abs(ctx.forces() - s.forces()).max() < 1e-13
    fn = s.numeric_forces(np.arange(3), delta=1e-5)
    assert np.abs(ctx.forces() - fn).max() < 1e-8


def test_forces_vs_finite_differences_gpu(workdir):
    say("water192 forces vs central differences of GPU energies")
    d, r, e, box = cases.water_case(workdir)
    m, ctx = _engine(d)
    ctx.set_structure(r, e, box)
    ctx.evaluate(0)
    f = ctx.forces()
    delta = 1e-4
    for i in (0, 1, 2, 150):
        for c in range(3):
            ep = []
            for sgn in (1, -1):
                rr = r.copy()
                rr[i, c] += sgn * delta
                rr = nio.remap(rr, box)
                ctx.set_positions(rr)
                ep.append(ctx.evaluate(-1))
            fn = -(ep[0] - ep[1]) / (2 * delta)
            assert abs(fn - f[i, c]) < 10 * delta ** 2
    ctx.set_positions(r)


def test_deterministic(workdir):
    say("determinism")
    d, r, e, box = cases.water_case(workdir)
    m, ctx = _engine(d)
    ctx.set_structure(r, e, box)
    e1 = ctx.evaluate(0)
    f1 = ctx.forces()
    e2 = ctx.evaluate(0)
    f2 = ctx.forces()
    assert e1 == e2 and np.array_equal(f1, f2)


def test_sf_statistics_gpu(workdir):
    say("SF statistics (nnp-scaling definition)")
    d, r, e, box = cases.water_case(workdir)
    m, ctx = _engine(d, scaling=False)
    ctx.set_structure(r, e, box)
    ctx.calculate_neighbor_list(m.max_cutoff)
    s = OStruct(OModel.from_dir(d, scaling=False), r, e, box)
    s.neighbor_list(m.max_cutoff, cells=True)
    for el in range(2):
        n, st = ctx.sf_statistics(el)
        on, ost = s.sf_statistics(el)
        assert n == on
        assert np.allclose(st, ost, rtol=1e-11, atol=1e-14)


def test_water24576_against_oracle(workdir):
    say("water 24,576 atoms: neighbour list bit-exact, E and F vs oracle")
    d, _, _, _ = cases.water_case(workdir)
    r, e, box = W.water(8192, 62.1, 99)
    m, ctx = _engine(d)
    ctx.set_structure(r, e, box)
    ctx.calculate_neighbor_list(m.max_cutoff)
    s = OStruct(OModel.from_dir(d), r, e, box)
    s.neighbor_list(m.max_cutoff, cells=True)
    _neighbors_equal(ctx, s, r, box)
    e_gpu = ctx.evaluate(0)
    e_cpu = s.evaluate(cells=True)
    assert abs(e_gpu - e_cpu) <= E_REL * abs(e_cpu)
    f = ctx.forces()
    assert np.abs(f - s.forces()).max() < F_ABS
    # translation invariance: forces sum to zero
    assert np.abs(f.sum(0)).max() < 1e-8
