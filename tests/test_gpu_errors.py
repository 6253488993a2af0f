"""Error paths of the C ABI on the GPU: a rejected load leaves the context
without a structure, so no later call launches kernels over stale sizes.

The reference throws from Structure / Mode on bad input
(ElementMap.cpp:89-104 for unknown elements, Mode.cpp:172-173); the C ABI
returns an error code and the context then holds nothing to evaluate."""
import sys

import numpy as np
import pytest

from n2p2_amd import engine
from tests import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def test_rejected_structure_leaves_no_stale_state(workdir):
    d, r, e, box = cases.water_case(workdir)
    m = engine.Model.load_dir(d)
    ctx = engine.Context(m)
    say("valid water 192, then element index out of range")
    ctx.set_structure(r, e, box)
    e0 = ctx.evaluate(engine.FORCES_REFERENCE)
    assert np.isfinite(e0)
    # A larger structure with one bad element index: rejected before any
    # state of the context changes, and the old structure is forgotten.
    r2 = np.concatenate([r, r[:50] + 0.1])
    e2 = np.concatenate([e, e[:50]]).astype(np.int32)
    e2[-1] = 7
    with pytest.raises(RuntimeError, match="Element index out of range"):
        ctx.set_structure(r2, e2, box)
    assert ctx.num_atoms() == 0
    with pytest.raises(RuntimeError, match="No structure loaded"):
        ctx.evaluate(engine.FORCES_REFERENCE)
    with pytest.raises(RuntimeError, match="No structure loaded"):
        ctx.calculate_symmetry_function_groups()
    # A bad batch and a bad caller-given list are rejected the same way.
    with pytest.raises(RuntimeError):
        ctx.set_structures([(r2, e2, box)])
    with pytest.raises(RuntimeError, match="No structure loaded"):
        ctx.evaluate(engine.FORCES_REFERENCE)
    off = np.array([1, 2], dtype=np.int64)
    with pytest.raises(RuntimeError, match=r"offsets\[0\] must be 0"):
        ctx.set_neighbor_list(np.zeros(1, np.int32), off,
                              np.zeros(2, np.int32), np.zeros(2, np.int32),
                              np.ones((2, 3)))
    with pytest.raises(RuntimeError, match="No structure loaded"):
        ctx.evaluate(engine.FORCES_REFERENCE)
    # The context is usable again after a valid load, with the same result.
    ctx.set_structure(r, e, box)
    assert ctx.evaluate(engine.FORCES_REFERENCE) == e0
    ctx.close()
