"""Compact (polynomial) symmetry functions, types 20-25 (SymGrpCompRad,
SymGrpCompAngn, SymGrpCompAngw and their weighted variants), on the device
against the oracle.

The oracle's compact functions are pinned by the reference's fixture
values (test/cpp/fixtures.h; checked on the GPU in test_gpu_parity.py), by
the asymmetric-core identity and by finite differences (test_oracle.py).
Here a B/C/N model has every compact family and core (p1-p4, p2a, e),
negative rl, angle intervals centred on 0 and 180 degrees and adjacent
members with equal (rl, rc) and different cores.

The compact pass runs in FP64 on both precision paths.  Tolerances as the
other FP64 parity tests: energy 1e-10 relative, forces 1e-9 eV/A, G and
dE/dG 1e-11 of their scale, per-slot pair forces 1e-10; the mixed path
against FP64 as the specialised-kernel tests (G 1e-5 of its scale, energy
1e-6 relative, forces 1e-4 eV/A).
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
    print("[gpu-compact]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ccase(tmp_path_factory):
    return cases.compact_case(tmp_path_factory.mktemp("compact"), natoms=96,
                              seed=7)


def test_compact_parity_fp64(ccase):
    d, r, e, box = ccase
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    s = OStruct(OModel.from_dir(d), r, e, box)
    for mode in (0, 1):
        say("compact 96 atoms, FP64, force mode", mode)
        en = ctx.evaluate(mode)
        oe = s.evaluate(cells=True, force_mode=mode)
        assert abs(en - oe) <= 1e-10 * abs(oe)
        G, oG = ctx.G(), s.G()
        gs = np.maximum(np.abs(oG).max(0), 1e-300)
        say("  G rel %.3g" % np.max(np.abs(G - oG) / gs))
        assert np.max(np.abs(G - oG) / gs) < 1e-11
        D, oD = ctx.dEdG(), s.dEdG()
        assert np.max(np.abs(D - oD)) / np.abs(oD).max() < 1e-11
        f, of = ctx.forces(), s.forces()
        say("  max |F - F_oracle| %.3g" % np.abs(f - of).max())
        assert np.abs(f - of).max() < 1e-9
    ctx.evaluate(0)
    s.evaluate(cells=True)
    sf, pf = ctx.pair_forces()
    osf, opf = s.pair_forces()
    say("  pair forces %.3g, self %.3g" % (np.abs(pf - opf).max(),
                                            np.abs(sf - osf).max()))
    assert np.abs(pf - opf).max() < 1e-10
    assert np.abs(sf - osf).max() < 1e-10


def test_compact_mixed_and_specialized(ccase):
    """Mixed precision and the specialised kernels with compact groups
    present (the compact pass is the same FP64 code on every path)."""
    d, r, e, box = ccase
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    out = {}
    for prec in (engine.PRECISION_FP64, engine.PRECISION_MIXED):
        ctx.set_precision(prec)
        for spec in (True, False):
            ctx.set_specialized(spec)
            en = ctx.evaluate(0)
            out[prec, spec] = (en, ctx.G(), ctx.forces())
    e0, G0, F0 = out[engine.PRECISION_FP64, False]
    gs = np.maximum(np.abs(G0).max(0), 1e-300)
    for key, (en, G, F) in out.items():
        mixed = key[0] == engine.PRECISION_MIXED
        tg, te, tf = (1e-5, 1e-6, 1e-4) if mixed else (1e-13, 1e-13, 1e-12)
        err = (np.max(np.abs(G - G0) / gs), abs(en - e0) / abs(e0),
               np.abs(F - F0).max())
        say("  precision %d specialised %d: G %.3g E %.3g F %.3g"
            % (key[0], key[1], *err))
        assert err[0] <= tg and err[1] <= te and err[2] <= tf


def test_compact_batch_and_statistics(tmp_path):
    """A batch of three structures equals one-at-a-time evaluation
    bitwise, and the raw SF statistics equal the oracle's."""
    d, r0, e0, b0 = cases.compact_case(tmp_path, natoms=40, seed=3)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    structs = [W.random_structure(n, seed=30 + n) for n in (19, 40, 27)]
    single = []
    for r, e, box in structs:
        ctx.set_structure(r, e, box)
        single.append((ctx.evaluate(0), ctx.forces()))
    ctx.set_structures(structs)
    ctx.evaluate(0)
    eb = ctx.structure_energies()
    fb = ctx.forces()
    o = 0
    for k, (r, e, box) in enumerate(structs):
        assert eb[k] == single[k][0]
        assert np.array_equal(fb[o:o + len(e)], single[k][1])
        o += len(e)
    mu = engine.Model.load_dir(d, scaling=False)
    cu = engine.Context(mu)
    cu.set_structure(r0, e0, b0)
    cu.calculate_neighbor_list(mu.max_cutoff)
    s = OStruct(OModel.from_dir(d, scaling=False), r0, e0, b0)
    s.neighbor_list(mu.max_cutoff, cells=True)
    for k in range(mu.num_elements):
        n, st = cu.sf_statistics(k)
        on, ost = s.sf_statistics(k)
        assert n == on
        assert np.allclose(st, ost, rtol=1e-11, atol=1e-14)




def test_compact_and_weighted_md_graph_equals_eager(tmp_path):
    """Graph-captured MD steps run the compact pass (and the weighted
    groups) like the eager steps, bitwise, over list rebuilds; the forces
    at the start equal a single-point evaluation to 1e-12."""
    for name, case in (("compact", cases.compact_case),
                       ("weighted", cases.weighted_case)):
        d, r, e, box = case(tmp_path / name, natoms=64, seed=11)
        masses = np.array([10.81, 12.011, 14.007])
        v = W.maxwell_boltzmann(masses[e], 300.0, seed=2)
        m = engine.Model.load_dir(d)
        c0 = engine.Context(m)
        c0.set_structure(r, e, box)
        c0.evaluate(engine.FORCES_CONSISTENT)
        f0 = c0.forces()
        out = []
        for graph in (True, False):
            c = engine.Context(m)
            c.set_structure(r, e, box)
            c.md_init(masses, 0.5, skin=0.05, v=v, graph=graph)
            if graph:
                # same kernels on the filtered skin list; the MD gather may
                # sum in another order
                assert np.abs(c.forces() - f0).max() <= 1e-12 * max(
                    1.0, np.abs(f0).max())
            ep, ek = c.md_run(20)
            rt, vt = c.md_state()
            out.append((ep, ek, rt, vt, c.md_rebuilds()))
        say("%s MD graph vs eager: rebuilds %d %d, epot %.6g -> %.6g"
            % (name, out[0][4], out[1][4], out[0][0][0], out[0][0][-1]))
        assert out[0][4] == out[1][4] and out[0][4] >= 1
        for a, b in zip(out[0][:4], out[1][:4]):
            assert np.array_equal(a, b)
