"""Velocity-Verlet MD on the device (hdnnp_md_*) with a Verlet-skin list.

- The skin list filtered to rc equals a fresh neighbour list bit for bit
  (indices, images, d, dr).
- After many steps with rebuilds, the MD forces and energy equal a fresh
  evaluation at the same positions: energy 1e-10 relative, forces 1e-9
  eV/A (only the summation order of the SF sums may differ).
- Ten steps agree with a velocity-Verlet loop on the oracle's FP64 forces:
  positions within 1e-10 A, energies within 1e-10 relative.
- The HIP-graph steps equal the same kernels issued one by one, bitwise,
  over rebuilds (each rebuild captures the graphs again).
- NVE energy: the total-energy error of velocity Verlet falls as dt^2
  (halving dt over the same time cuts it about fourfold), for FP64 and for
  the mixed path, whose error stays within twice the FP64 one.
"""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from n2p2_amd import workloads as W
from tests import cases
from tests.oracle_lib import OModel, OStruct

pytestmark = pytest.mark.gpu

DT = 0.25


def say(*a):
    print("[gpu-md]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def water(tmp_path_factory):
    d, r, e, box = cases.water_case(tmp_path_factory.mktemp("md"))
    masses = np.array(W.WATER_MASSES)
    v = W.maxwell_boltzmann(masses[e], 300.0, seed=5)
    return d, r, e, box, masses, v


def _ctx(d, prec=engine.PRECISION_FP64):
    m = engine.Model.load_dir(d)
    c = engine.Context(m)
    c.set_precision(prec)
    return m, c


def test_md_filtered_list_equals_fresh(water):
    d, r, e, box, masses, v = water
    m, c = _ctx(d)
    c.set_structure(r, e, box)
    c.md_init(masses, DT, skin=1.0, v=v)
    m2, f = _ctx(d)
    f.set_structure(r, e, box)
    f.calculate_neighbor_list(m.max_cutoff)
    a, b = c.neighbors(), f.neighbors()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    say("filtered skin list == fresh list:", len(a[1]), "entries")


def test_md_forces_equal_fresh_evaluation(water):
    d, r, e, box, masses, v = water
    m, c = _ctx(d)
    c.set_structure(r, e, box)
    c.md_init(masses, DT, skin=0.05, v=v)
    ep, ek = c.md_run(40)
    nreb = c.md_rebuilds()
    rt, vt = c.md_state()
    F = c.forces()
    m2, f = _ctx(d)
    f.set_structure(W.remap(rt, box), e, box)
    en = f.evaluate(engine.FORCES_CONSISTENT)
    ferr = float(np.abs(F - f.forces()).max())
    erel = abs(ep[-1] - en) / abs(en)
    say("40 steps, %d rebuilds: e_rel %.2g, f %.2g" % (nreb, erel, ferr))
    assert nreb > 2
    assert c.md_partner_misses() == 0
    assert erel <= 1e-10 and ferr <= 1e-9


def test_md_vs_oracle_verlet(water):
    d, r, e, box, masses, v = water
    m, c = _ctx(d)
    c.set_structure(r, e, box)
    c.md_init(masses, DT, skin=1.0, v=v)
    r0, v0 = c.md_state()
    ep, ek = c.md_run(10)
    rt, vt = c.md_state()
    om = OModel.from_dir(d)
    acc = W.FTOA / masses[e][:, None]

    def forces(x):
        s = OStruct(om, x, e, box)
        en = s.evaluate(cells=True, force_mode=1)
        return en, s.forces()

    x, u = r0.copy(), v0.copy()
    _, F = forces(x)
    oep = []
    for _ in range(10):
        u = u + 0.5 * DT * acc * F
        x = x + DT * u
        en, F = forces(x)
        u = u + 0.5 * DT * acc * F
        oep.append(en)
    rerr = float(np.abs(rt - x).max())
    verr = float(np.abs(vt - u).max())
    eerr = float(np.max(np.abs(ep - oep) / np.abs(oep)))
    say("10 steps vs oracle Verlet: r %.2g A, v %.2g A/fs, e_rel %.2g"
        % (rerr, verr, eerr))
    assert rerr <= 1e-10 and verr <= 1e-10 and eerr <= 1e-10
    ekin = 0.5 * (masses[e][:, None] * vt ** 2).sum() / W.FTOA
    assert abs(ek[-1] - ekin) <= 1e-10 * ekin


def test_md_graph_equals_eager(water):
    d, r, e, box, masses, v = water
    out = []
    for graph in (True, False):
        m, c = _ctx(d)
        c.set_structure(r, e, box)
        c.md_init(masses, DT, skin=0.05, v=v, graph=graph)
        ep, ek = c.md_run(30)
        ep2, ek2 = c.md_run(15)
        rt, vt = c.md_state()
        out.append((np.concatenate([ep, ep2]), np.concatenate([ek, ek2]),
                    rt, vt, c.md_rebuilds()))
    say("graph vs eager: rebuilds", out[0][4], out[1][4])
    assert out[0][4] == out[1][4] and out[0][4] > 2
    for a, b in zip(out[0][:4], out[1][:4]):
        assert np.array_equal(a, b)


def _spread(d, r, e, box, masses, v, prec, dt, nsteps):
    m, c = _ctx(d, prec)
    c.set_structure(r, e, box)
    c.md_init(masses, dt, skin=1.0, v=v)
    ep, ek = c.md_run(nsteps)
    et = ep + ek
    n = len(e)
    return (et.max() - et.min()) / n, abs(et[-1] - et[0]) / n, ek.mean()


def test_md_energy_conservation(water):
    """Velocity Verlet's energy error falls as dt^2: halving dt over the
    same 50 fs cuts the total-energy spread about fourfold (asserted below
    0.4).  That holds only if the forces are the gradient of the energy.
    The mixed path's spread stays within twice the FP64 one."""
    d, r, e, box, masses, v = water
    res = {}
    for prec in (engine.PRECISION_FP64, engine.PRECISION_MIXED):
        for dt, ns in ((DT, 200), (DT / 2, 400)):
            res[prec, dt] = _spread(d, r, e, box, masses, v, prec, dt, ns)
            say("precision %d dt %.3f fs: spread %.3g eV/atom, drift %.3g "
                "eV/atom, <ekin> %.3f eV" % ((prec, dt) + res[prec, dt]))
    for prec in (engine.PRECISION_FP64, engine.PRECISION_MIXED):
        assert res[prec, DT / 2][0] <= 0.4 * res[prec, DT][0]
    assert res[engine.PRECISION_MIXED, DT][0] <= \
        2.0 * res[engine.PRECISION_FP64, DT][0]
