"""Batched structures (hdnnp_set_structures) against the oracle evaluated
one structure at a time, the way nnp-dataset.cpp:206-211 loops
Mode::evaluateNNP over a data set.

Bars (plain FP64): every structure's neighbour list bit-exact (offsets,
index, d, dr) with its own image counts; structure energies within 1e-10
relative; forces within 1e-9 eV/A; G and dE/dG within 1e-11 of their
scale.  Mixed precision: the mixed-path bounds against the FP64 oracle
(energy 1e-6 relative, force RMSE 1e-4 and largest error 1e-3 eV/A).

The batch mixes what the single-structure path meets separately: 16- and
17-atom cells shorter than the cutoff (two images per direction, the atom's
own images in its list), a sheared (triclinic) cell, an open cluster and
larger cells, in one launch per kernel.
"""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import workloads as W
from tests import cases
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu

E_REL = 1e-10
F_ABS = 1e-9
G_REL = 1e-11
MIXED = dict(e_rel=1e-6, f_rmse=1e-4, f_max=1e-3)


def say(*a):
    print("[gpu-batch]", *a, flush=True)
    sys.stdout.flush()


def _structures():
    out = []
    for k, n in enumerate((16, 17, 23, 64, 150, 41)):
        out.append(W.random_structure(n, seed=100 + k))
    r, e, box = W.random_structure(48, seed=120)
    r, box = W.sheared(r, box, [[1.0, 0.0, 0.0], [0.3, 1.0, 0.0],
                                [0.2, -0.25, 1.0]])
    out.append((r, e, box))
    r, e, _ = W.random_structure(30, seed=121)
    out.append((r, e, None))
    return out


@pytest.fixture(scope="module")
def batch(tmp_path_factory):
    d, _, _, _ = cases.mixed_case(tmp_path_factory.mktemp("batch"), natoms=96,
                                  seed=3)
    om = OModel.from_dir(d)
    sts = _structures()
    oracle = []
    for r, e, box in sts:
        s = OStruct(om, r, e, box)
        oracle.append(s)
    return d, sts, oracle


def _offsets(sts):
    return np.cumsum([0] + [len(e) for _, e, _ in sts])


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_fp64_vs_oracle(batch, mode):
    d, sts, oracle = batch
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structures(sts)
    say("batch of", len(sts), "structures,", ctx.n, "atoms, force mode", mode)
    assert ctx.num_structures() == len(sts)
    etot = ctx.evaluate(mode)
    es = ctx.structure_energies()
    pbc = ctx.structure_pbc()
    off, idx, img, dd, dr = ctx.neighbors()
    G, D, F = ctx.G(), ctx.dEdG(), ctx.forces()
    ao = _offsets(sts)
    assert abs(etot - es.sum()) <= 1e-12 * abs(etot)
    for k, s in enumerate(oracle):
        a0, a1 = ao[k], ao[k + 1]
        oe = s.evaluate(cells=False, force_mode=mode)
        ooff, oidx, od, odr = s.neighbors()
        assert np.array_equal(pbc[k], s.pbc()), (k, pbc[k], s.pbc())
        q0, q1 = off[a0], off[a1]
        assert np.array_equal(off[a0:a1 + 1] - q0, ooff), k
        assert np.array_equal(idx[q0:q1] - a0, oidx), k
        assert np.array_equal(dd[q0:q1], od), k
        assert np.array_equal(dr[q0:q1], odr), k
        erel = abs(es[k] - oe) / abs(oe)
        ferr = float(np.abs(F[a0:a1] - s.forces()).max())
        oG = s.G()
        gs = np.maximum(np.abs(oG).max(0), 1e-300)
        gerr = float(np.max(np.abs(G[a0:a1, :oG.shape[1]] - oG) / gs))
        oD = s.dEdG()
        derr = float(np.max(np.abs(D[a0:a1, :oD.shape[1]] - oD))
                     / max(np.abs(oD).max(), 1e-300))
        say("  structure %d: %d atoms, pbc %s, %d entries, e_rel %.2g, "
            "f %.2g, G %.2g, dEdG %.2g" % (k, a1 - a0, pbc[k].tolist(),
                                           q1 - q0, erel, ferr, gerr, derr))
        assert erel <= E_REL
        assert ferr <= F_ABS
        assert gerr <= G_REL and derr <= G_REL


def test_batch_mixed_vs_oracle(batch):
    d, sts, oracle = batch
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_precision(engine.PRECISION_MIXED)
    ctx.set_structures(sts)
    ctx.evaluate(0)
    es, F = ctx.structure_energies(), ctx.forces()
    ao = _offsets(sts)
    for k, s in enumerate(oracle):
        oe = s.evaluate(cells=False, force_mode=0)
        df = F[ao[k]:ao[k + 1]] - s.forces()
        err = dict(e_rel=abs(es[k] - oe) / abs(oe),
                   f_rmse=float(np.sqrt(np.mean(df ** 2))),
                   f_max=float(np.abs(df).max()))
        say("  mixed structure %d:" % k, {a: "%.3g" % v for a, v in err.items()})
        for a, b in MIXED.items():
            assert err[a] <= b, (k, a, err[a], b)


def test_batch_equals_single(batch):
    """A batch gives every structure exactly what hdnnp_set_structure gives
    it alone (same kernels, same per-atom order): bitwise in FP64."""
    d, sts, _ = batch
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structures(sts)
    ctx.evaluate(0)
    es, F = ctx.structure_energies(), ctx.forces()
    ao = _offsets(sts)
    one = engine.Context(m)
    for k, (r, e, box) in enumerate(sts):
        one.set_structure(r, e, box)
        en = one.evaluate(0)
        assert np.array_equal(one.forces(), F[ao[k]:ao[k + 1]]), k
        assert abs(en - es[k]) <= 1e-13 * abs(en), (k, en, es[k])


def test_batch_then_single_then_batch(batch):
    """Switching between a batch and a single structure on one context
    rebuilds every per-layout buffer (atom order, offsets, image counts)."""
    d, sts, oracle = batch
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structures(sts[:3])
    ctx.evaluate(0)
    first = ctx.structure_energies().copy()
    r, e, box = sts[4]
    ctx.set_structure(r, e, box)
    en = ctx.evaluate(0)
    oe = oracle[4].evaluate(cells=False, force_mode=0)
    assert abs(en - oe) <= E_REL * abs(oe)
    assert np.array_equal(ctx.structure_energies(), [en])
    ctx.set_structures(sts[:3])
    ctx.evaluate(0)
    assert np.array_equal(ctx.structure_energies(), first)
    with pytest.raises(RuntimeError, match="per structure"):
        ctx.pbc()


def test_batch_c2_full(batch):
    """The full batch: 512 structures of 16-512 atoms in one evaluation.
    Six structures (smallest, largest, four more) against the oracle at the
    FP64 bars; the batch total equals the sum of the structure energies;
    the mixed path's force RMSE over all atoms against the FP64 path."""
    d, _, _ = batch
    sts = W.batch_structures(512, 16, 512, seed=2024)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structures(sts)
    say("C2 batch:", len(sts), "structures,", ctx.n, "atoms")
    etot = ctx.evaluate(0)
    es, F = ctx.structure_energies(), ctx.forces()
    assert abs(etot - es.sum()) <= 1e-12 * abs(etot)
    ao = _offsets(sts)
    sizes = np.diff(ao)
    rng = np.random.default_rng(7)
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    pick |= set(int(x) for x in rng.choice(len(sts), 4, replace=False))
    om = OModel.from_dir(d)
    for k in sorted(pick):
        r, e, box = sts[k]
        s = OStruct(om, r, e, box)
        oe = s.evaluate(cells=True, force_mode=0)
        erel = abs(es[k] - oe) / abs(oe)
        ferr = float(np.abs(F[ao[k]:ao[k + 1]] - s.forces()).max())
        say("  structure %d (%d atoms): e_rel %.2g, f %.2g" % (k, sizes[k],
                                                              erel, ferr))
        assert erel <= E_REL and ferr <= F_ABS
    # The mixed forms against FP64 over all atoms: FP32 SFs with FP32, FP16
    # and BF16 network layers (split 16-bit operands, FP32 accumulation),
    # each within the 1 meV/A force RMSE bound.
    # BF16's 16-bit split operands leave about 1e-5 relative in every
    # atomic energy, and those add up within a structure.
    for name, prec, e_rel in (("FP32", engine.PRECISION_MIXED, 1e-6),
                              ("FP16", engine.PRECISION_MIXED16, 1e-6),
                              ("BF16", engine.PRECISION_MIXED_BF16, 1e-4)):
        ctx.set_precision(prec)
        ctx.evaluate(0)
        df = ctx.forces() - F
        rmse = float(np.sqrt(np.mean(df ** 2)))
        de = float(np.max(np.abs(ctx.structure_energies() - es) / np.abs(es)))
        say("  mixed (%s networks) vs FP64 over the batch: force RMSE %.3g, "
            "max %.3g, energy rel %.3g" % (name, rmse, np.abs(df).max(), de))
        assert rmse <= MIXED["f_rmse"] and np.abs(df).max() <= MIXED["f_max"]
        assert de <= e_rel
