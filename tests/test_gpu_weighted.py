"""Element-weighted symmetry functions, types 12 (SymFncExpRadWeighted,
SymGrpExpRadWeighted.cpp:144-252) and 13 (SymFncExpAngnWeighted,
SymGrpExpAngnWeighted.cpp:166-385), on the device against the oracle.

The oracle's weighted functions are pinned by the reference's fixture
values (test/cpp/fixtures.h, Mono / Dual / Mixed; test_oracle.py) and by
finite differences; the fixture values themselves are checked on the GPU in
test_gpu_parity.py.  Here a B/C/N model mixes types 2, 3, 12 and 13 with
two cutoff radii, shifted Gaussians and non-integer zeta.

Tolerances: plain FP64 as the other parity tests (energy 1e-10 relative,
forces 1e-9 eV/A, G and dE/dG 1e-11 of their scale, per-slot pair forces
1e-10); mixed precision against FP64 as the specialised-kernel tests (G
1e-5 of its scale, energy 1e-6 relative, forces 1e-4 eV/A).
"""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from tests import cases
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-weighted]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def wcase(tmp_path_factory):
    return cases.weighted_case(tmp_path_factory.mktemp("weighted"), natoms=96,
                               seed=5)


def test_weighted_parity_fp64(wcase):
    d, r, e, box = wcase
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    s = OStruct(OModel.from_dir(d), r, e, box)
    for mode in (0, 1):
        say("weighted 96 atoms, FP64, force mode", mode)
        en = ctx.evaluate(mode)
        oe = s.evaluate(cells=True, force_mode=mode)
        assert abs(en - oe) <= 1e-10 * abs(oe)
        G, oG = ctx.G(), s.G()
        gs = np.maximum(np.abs(oG).max(0), 1e-300)
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
    assert np.abs(pf - opf).max() < 1e-10
    assert np.abs(sf - osf).max() < 1e-10


def test_weighted_specialized_and_mixed(wcase):
    """Specialised (run-time compiled) against generic kernels, FP64 and
    mixed; mixed against FP64."""
    d, r, e, box = wcase
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_structure(r, e, box)
    out = {}
    for prec in (engine.PRECISION_FP64, engine.PRECISION_MIXED):
        ctx.set_precision(prec)
        for spec in (True, False):
            ctx.set_specialized(spec)
            assert ctx.specialized_active() == spec
            en = ctx.evaluate(0)
            out[prec, spec] = (en, ctx.G(), ctx.forces())
    ctx.set_specialized(True)
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


def test_weighted_batch_equals_single(tmp_path):
    """Three structures of different sizes in one batch against one at a
    time (bitwise on the FP64 path)."""
    from n2p2_amd import workloads as W
    d, _, _, _ = cases.weighted_case(tmp_path, natoms=40, seed=3)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    structs = [W.random_structure(n, seed=20 + n) for n in (17, 40, 29)]
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
