"""The allocation-free scatter (hdnnp_get_scattered_forces_noalloc) against
hdnnp_get_scattered_forces, through the C ABI: the same kernel staged in the
library's fixed device array instead of a context buffer, so the forces
must be equal bit for bit.  Two calls in a row on one context, the second
with a smaller caller-given list (fewer centres and targets), and a list
whose targets exceed the array's capacity, which is refused."""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from tests import cases
from tests.lammps_like import full_lists, ghost_domain

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)
    sys.stdout.flush()


def test_noalloc_scatter_equals_scatter(workdir):
    d, r, e, box = cases.water_case(workdir)
    m = engine.Model.load_dir(d)
    x, el, owner, nl = ghost_domain(r, e, box, m.max_cutoff)
    off, idx, dr = full_lists(x, nl, m.max_cutoff)
    ctx = engine.Context(m)
    # First call: all 192 centres, 192 local + ghost targets.
    ctx.set_neighbor_list(el[:nl], off, idx, el[idx], dr)
    ctx.evaluate(engine.FORCES_CONSISTENT)
    nout = len(x)
    a = ctx.scattered_forces_noalloc(nout)
    b = ctx.scattered_forces(nout)
    say("192 centres, %d targets: max |F| %.3g" % (nout, np.abs(a).max()))
    assert np.abs(a).max() > 0.0
    assert np.array_equal(a, b)
    # Second call, smaller: the first 64 centres and only their targets.
    k = 64
    sub = idx[:off[k]]
    nout2 = int(max(sub.max(), k - 1)) + 1
    ctx.set_neighbor_list(el[:k], off[:k + 1], sub, el[sub], dr[:off[k]])
    ctx.evaluate(engine.FORCES_CONSISTENT)
    a2 = ctx.scattered_forces_noalloc(nout2)
    b2 = ctx.scattered_forces(nout2)
    say("64 centres, %d targets" % nout2)
    assert nout2 < nout
    assert np.abs(a2).max() > 0.0
    assert np.array_equal(a2, b2)
    # A target index beyond the array's capacity (524,288 atoms) is refused
    # before anything runs.
    big = idx[:off[k]].copy()
    big[0] = 600000
    ctx.set_neighbor_list(el[:k], off[:k + 1], big, el[sub], dr[:off[k]])
    ctx.evaluate(engine.FORCES_CONSISTENT)
    with pytest.raises(RuntimeError, match="beyond the staging"):
        ctx.scattered_forces_noalloc(600001)
    ctx.close()
