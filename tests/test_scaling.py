"""nnp-scaling (n2p2_amd/scaling.py), CPU side.

Golden: tests/golden/nnp_scaling_H2O/ holds the reference's own example
data set examples/nnp-train/H2O_RPBE-D3 (input.nn, 20 structures of
input.data, and the scaling.data nnp-scaling wrote for them).

- The header writer reproduces the example file's header byte for byte.
- The multi-structure reader finds the 20 structures.
- The oracle's SF statistics over the 20 structures (normalised units, as
  nnp-scaling converts them) reproduce the example's scaling.data: min, max
  and mean within 1e-14 relative, sigma within 1e-12 (the oracle's
  per-structure statistics are combined here through mean and sigma).
"""
import os
import shutil

import numpy as np

from n2p2_amd import io as nio
from n2p2_amd import scaling as SC
from tests.oracle_lib import OModel, OStruct

GOLD = os.path.join(os.path.dirname(__file__), "golden", "nnp_scaling_H2O")


def _golden_rows():
    return np.array([[float(x) for x in l.split()]
                     for l in open(os.path.join(GOLD, "scaling.data"))
                     if l.strip() and not l.startswith("#")])


def test_header_matches_reference_file():
    size, name, info = SC.SCALING_COLUMNS
    ours = SC.file_header(["Symmetry function scaling data."], size, name,
                          info)
    theirs = [l for l in open(os.path.join(GOLD, "scaling.data"))
              if l.startswith("#")]
    assert ours == theirs


def test_writer_round_trip(tmp_path):
    rows = _golden_rows()
    stats = [rows[rows[:, 0] == e][:, 2:] for e in (1, 2)]
    p = str(tmp_path / "scaling.data")
    SC.write_scaling_file(p, stats)
    assert open(p).read() == open(os.path.join(GOLD, "scaling.data")).read()


def test_read_structures():
    sts = nio.read_structures(os.path.join(GOLD, "input.data"),
                              lambda s: {"H": 0, "O": 1}.get(s, -1))
    assert len(sts) == 20
    for r, e, box, energy in sts:
        assert box is not None and r.shape == (len(e), 3)


def test_oracle_statistics_match_reference_scaling(tmp_path):
    d = str(tmp_path)
    shutil.copy(os.path.join(GOLD, "input.nn"), d)
    for z in (1, 8):
        with open(os.path.join(d, "weights.%03d.data" % z), "w") as f:
            f.write("0.0\n" * 5000)
    om = OModel(os.path.join(d, "input.nn"), None,
                os.path.join(d, "weights.%03zu.data"))
    sts = nio.read_structures(os.path.join(GOLD, "input.data"),
                              lambda s: {"H": 0, "O": 1}.get(s, -1))
    per = {e: [] for e in range(om.num_elements)}
    for r, e, box, _ in sts:
        s = OStruct(om, r, e, box)
        s.to_normalized()
        s.neighbor_list(om.max_cutoff)
        for k in range(om.num_elements):
            per[k].append(s.sf_statistics(k))
    rows = _golden_rows()
    for k in range(om.num_elements):
        ref = rows[rows[:, 0] == k + 1][:, 2:]
        N = sum(n for n, _ in per[k])
        mn = np.min([st[:, 0] for n, st in per[k] if n], axis=0)
        mx = np.max([st[:, 1] for n, st in per[k] if n], axis=0)
        s1 = np.sum([st[:, 2] * n for n, st in per[k]], axis=0)
        s2 = np.sum([st[:, 3] ** 2 * (n - 1) + (st[:, 2] * n) ** 2 / n
                     for n, st in per[k] if n], axis=0)
        mean = s1 / N
        sig = np.sqrt((s2 - s1 * s1 / N) / (N - 1))
        rel = lambda a, b: float(np.max(np.abs(a - b) / np.abs(b)))
        assert rel(mn, ref[:, 0]) <= 1e-14
        assert rel(mx, ref[:, 1]) <= 1e-14
        assert rel(mean, ref[:, 2]) <= 1e-14
        assert rel(sig, ref[:, 3]) <= 1e-12
