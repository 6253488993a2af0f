"""C ABI checks that need no GPU: the library loads, exports every function
include/hdnnp.h declares, and its host-side model setup (settings parsing,
SF sort, scaling) agrees with the oracle's."""
import os
import re

import numpy as np
import pytest

from n2p2_amd import engine
from tests import cases
from tests.oracle_lib import OModel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "hdnnp.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hdnnp_[A-Za-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = engine.load_library()
    decl = _declared()
    assert len(decl) >= 40
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(engine.EXPORTED) == decl


@pytest.mark.parametrize("name", cases.EXAMPLE_NAMES)
def test_host_model_setup_matches_oracle(name):
    d = os.path.join(cases.EXAMPLES, name)
    m = engine.Model.load_dir(d)
    o = OModel.from_dir(d)
    assert m.num_elements == o.num_elements
    assert m.max_cutoff == o.max_cutoff
    assert m.normalization() == o.normalization()
    for e in range(m.num_elements):
        a, b = m.sf_table(e), o.sf_table(e)
        assert a.shape == b.shape
        assert np.array_equal(a, b)


def test_host_model_setup_weighted_types(tmp_path):
    """Types 12 and 13 (no neighbour elements, relevant to every element)
    sort, group and index like the oracle's."""
    d, _, _, _ = cases.weighted_case(tmp_path)
    m = engine.Model.load_dir(d)
    o = OModel.from_dir(d)
    for e in range(m.num_elements):
        a, b = m.sf_table(e), o.sf_table(e)
        assert a.shape == b.shape == (41, 12)
        assert np.array_equal(a, b)
        assert set(a[:, 0].astype(int)) == {2, 3, 12, 13}


def test_host_model_setup_compact_types(tmp_path):
    """The compact types 20-25 (subtype, interval and angle ordering,
    grouping by centre and neighbour elements) sort like the oracle's."""
    d, _, _, _ = cases.compact_case(tmp_path)
    m = engine.Model.load_dir(d)
    o = OModel.from_dir(d)
    for e in range(m.num_elements):
        a, b = m.sf_table(e), o.sf_table(e)
        assert a.shape == b.shape == (54, 12)
        assert np.array_equal(a, b)


def test_model_errors_are_reported(tmp_path):
    with pytest.raises(RuntimeError, match="Could not open file"):
        engine.Model.load(str(tmp_path / "missing.nn"), None, "w.%03zu.data")
    p = tmp_path / "input.nn"
    p.write_text("number_of_elements 1\nelements H\ncutoff_type 1\n"
                 "symfunction_short H 13 0.01 -1.0 1.0 1.0 5.0\n")
    with pytest.raises(RuntimeError, match="outside this engine"):
        engine.Model.load(str(p), None, str(tmp_path / "w.%03zu.data"))
    p.write_text("number_of_elements 1\nelements H\ncutoff_type 1\n"
                 + "".join("symfunction_short H 20 H 0.0 %.1f p2\n"
                           % (3.0 + 0.05 * k) for k in range(97)))
    with pytest.raises(RuntimeError, match="outside this engine"):
        engine.Model.load(str(p), None, str(tmp_path / "w.%03zu.data"))
    p.write_text("number_of_elements 1\nelements H\ncutoff_type 1\n"
                 "symfunction_short H 5 H 0.0 5.0\n")
    with pytest.raises(RuntimeError, match="outside this engine"):
        engine.Model.load(str(p), None, str(tmp_path / "w.%03zu.data"))


def test_context_without_device_fails_cleanly():
    import ctypes
    lib = engine.load_library()
    if lib.hdnnp_ctx_create is None:
        pytest.skip("no symbol")
    d = os.path.join(cases.EXAMPLES, "Cu2S_PBE")
    m = engine.Model.load_dir(d)
    try:
        c = engine.Context(m, device=999)
    except RuntimeError as e:
        assert "ERROR" in str(e)
    else:  # pragma: no cover - there is no device 999
        c.close()
        pytest.fail("device 999 accepted")


def test_bindings_marshal_arguments_without_a_context():
    """The ctypes wrappers of the tuning and allocation-free entry points
    pass their arrays with the declared pointer types: called on a null
    context, each reaches the library and comes back with its null-argument
    error (no device is touched), not a ctypes argument error."""
    engine.load_library()
    c = object.__new__(engine.Context)
    c._h = None
    for call in (lambda: c.tuning(),
                 lambda: c.set_tuning(8, -1, -1, 1),
                 lambda: c.autotune(sweep=True),
                 lambda: c.autotune(),
                 lambda: c.scattered_forces_noalloc(4),
                 lambda: c.kernel_name("force")):
        with pytest.raises(RuntimeError, match="ERROR"):
            call()
