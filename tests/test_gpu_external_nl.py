"""Caller-given neighbour lists (the InterfaceLammps path) against the
oracle's periodic evaluation of the same box: a LAMMPS-like domain with
explicit ghost atoms (tests/lammps_like.py) evaluated through
hdnnp_set_neighbor_list, its forces scattered onto local and ghost atoms on
the device (InterfaceLammps::getForces, InterfaceLammps.cpp:533-621), must
give the oracle's energy (1e-10 relative), atomic energies and, after
folding the ghost forces onto their owners, the oracle's forces with
self-images (force mode 1, -dE/dr) within 1e-9 eV/A.  Both precisions; the
mixed path within its own bounds (1e-6 relative, 1e-5 eV/A).  The device
scatter is also checked against the reference's loop, restated on the host
over the per-slot forces: bit for bit."""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from tests import cases
from tests.lammps_like import full_lists, ghost_domain
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-extnl]", *a, flush=True)
    sys.stdout.flush()


def _loop_scatter(nout, off, idx, sf, pf):
    """InterfaceLammps::getForces (InterfaceLammps.cpp:546-581) as the
    reference loops: per centre its self force, then its pair forces onto
    the neighbours in list order."""
    f = np.zeros((nout, 3))
    for i in range(len(off) - 1):
        f[i] -= -sf[i]
        for q in range(off[i], off[i + 1]):
            f[idx[q]] -= -pf[q]
    return f


def _compare(d, r, e, box, prec, e_rel, f_abs):
    m = engine.Model.load_dir(d)
    s = OStruct(OModel.from_dir(d), r, e, box)
    ep = s.evaluate(cells=True, force_mode=1)
    fp = s.forces()
    rc = m.max_cutoff
    x, el, owner, nl = ghost_domain(r, e, box, rc)
    off, idx, dr = full_lists(x, nl, rc)
    say("  %d local, %d ghost atoms, %d entries, precision %d"
        % (nl, len(x) - nl, len(idx), prec))
    ext = engine.Context(m)
    ext.set_precision(prec)
    ext.set_neighbor_list(el[:nl], off, idx, el[idx], dr)
    ee = ext.evaluate(engine.FORCES_CONSISTENT)
    fg = ext.scattered_forces(len(x))
    # The device scatter is the reference's loop over the per-slot forces.
    sf, pf = ext.pair_forces()
    assert np.array_equal(fg, _loop_scatter(len(x), off, idx, sf, pf))
    folded = np.zeros((nl, 3))
    np.add.at(folded, owner, fg)
    de = abs(ee - ep) / abs(ep)
    df = float(np.abs(folded - fp).max())
    say("  energy rel %.3g, folded force max %.3g (oracle)" % (de, df))
    assert de <= e_rel
    assert df <= f_abs
    # The per-atom energies are the oracle's periodic ones (to the path's
    # relative bound on their scale).
    ea, oa = ext.atomic_energies(), s.atomic_energies()
    assert np.abs(ea - oa).max() <= 10 * e_rel * np.abs(oa).max()


@pytest.mark.parametrize("prec,e_rel,f_abs",
                         [(engine.PRECISION_FP64, 1e-10, 1e-9),
                          (engine.PRECISION_MIXED, 1e-6, 1e-5)])
def test_external_list_water192(workdir, prec, e_rel, f_abs):
    say("water192 as a LAMMPS-like domain")
    d, r, e, box = cases.water_case(workdir)
    _compare(d, r, e, box, prec, e_rel, f_abs)


def test_external_list_g9_small_box(workdir):
    """G2 + G9, POLY2, softplus in a cell smaller than twice the cutoff:
    atoms see their own periodic images as ghosts."""
    say("mixed 96 atoms (G2 + G9) as a LAMMPS-like domain")
    d, r, e, box = cases.mixed_case(workdir, natoms=96, seed=3)
    _compare(d, r, e, box, engine.PRECISION_FP64, 1e-10, 1e-9)


@pytest.mark.parametrize("which", ["weighted", "compact"])
def test_external_list_weighted_and_compact(workdir, which):
    """The element-weighted types 12, 13 and the compact types 20-25 through
    a LAMMPS-like domain: the compact pass adds its slot contributions
    before the scatter to ghosts."""
    say(which, "96 atoms as a LAMMPS-like domain")
    case = cases.weighted_case if which == "weighted" else cases.compact_case
    d, r, e, box = case(workdir, natoms=96, seed=5)
    _compare(d, r, e, box, engine.PRECISION_FP64, 1e-10, 1e-9)
    _compare(d, r, e, box, engine.PRECISION_MIXED, 1e-6, 1e-5)


def test_external_list_errors(workdir):
    d, r, e, box = cases.water_case(workdir)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    ctx.set_neighbor_list([0, 1], [0, 1, 2], [1, 0], [1, 0],
                          [[1.0, 0, 0], [-1.0, 0, 0]])
    with pytest.raises(RuntimeError, match="given by the caller"):
        ctx.calculate_neighbor_list(5.0)
    with pytest.raises(RuntimeError, match="scattered"):
        ctx.evaluate(0)
        ctx.forces()
    with pytest.raises(RuntimeError, match="out of range"):
        ctx.scattered_forces(1)


@pytest.mark.parametrize("case", ["water", "g9"])
def test_interface_lammps_driver(workdir, tmp_path, case):
    """nnp::InterfaceLammps driven as pair_style hdnnp drives it
    (n2p2_amd/nnp/bin/lammps_like): energy with offsets and folded forces in
    physical units against the periodic evaluation."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "n2p2_amd", "nnp", "bin", "lammps_like")
    assert os.path.exists(exe), "lammps_like not built"
    if case == "water":
        d, r, e, box = cases.water_case(workdir)
    else:
        d, r, e, box = cases.mixed_case(workdir, natoms=96, seed=3)
    say("InterfaceLammps driver:", case)
    m = engine.Model.load_dir(d)
    norm, mean_e, conv_e, conv_l = m.normalization()
    assert not norm      # the synthetic models are in physical units
    per = engine.Context(m)
    per.set_structure(r, e, box)
    ep = per.evaluate(engine.FORCES_CONSISTENT)
    ep += sum(m.energy_offset(int(k)) for k in e)
    fp = per.forces()
    rc = m.max_cutoff
    x, el, owner, nl = ghost_domain(r, e, box, rc)
    dom = tmp_path / "domain.txt"
    with open(dom, "w") as f:
        f.write("%d %d %d %.17g\n" % (m.num_elements, nl, len(x), rc * 1.001))
        for k in range(len(x)):
            f.write("%d %.17g %.17g %.17g\n" % (el[k] + 1, *x[k]))
    out = subprocess.run([exe, d, str(dom)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = out.stdout.splitlines()
    ee = float(lines[0].split()[1])
    fg = np.array([[float(v) for v in l.split()[2:5]] for l in lines[1:]])
    folded = np.zeros((nl, 3))
    np.add.at(folded, owner, fg)
    de = abs(ee - ep) / abs(ep)
    df = float(np.abs(folded - fp).max())
    say("  energy rel %.3g, folded force max %.3g" % (de, df))
    assert de <= 1e-10
    assert df <= 1e-9
