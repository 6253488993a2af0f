"""Shared test cases: model directories and structures, built from seeds.

Scaling files for the synthetic models are computed by the oracle (the
nnp-scaling definition: unscaled SFs over the structure, Dataset.cpp:1074),
so that engine and oracle read the same scaling.data.
"""
import json
import os

import numpy as np

from n2p2_amd import workloads as W
from tests.oracle_lib import OModel, OStruct

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
EXAMPLES = os.path.join(GOLDEN, "examples")
# The 2G examples of examples/nnp-predict with ExamplePrediction.h goldens.
# The last three use the compact SF types 20-22 in groups of 54 members.
EXAMPLE_NAMES = ["H2O_RPBE-D3", "Cu2S_PBE", "Anisole_SCAN", "DMABN_SCAN",
                 "Ethylbenzene_SCAN"]


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def scale_from_oracle(d, r, elem, box):
    """Writes d/scaling.data from the oracle's unscaled SF statistics."""
    m = OModel.from_dir(d, scaling=False)
    s = OStruct(m, r, elem, box)
    s.neighbor_list(m.max_cutoff, cells=True)
    stats = []
    for e in range(m.num_elements):
        n, st = s.sf_statistics(e)
        # Keep min < max so that min/max scaling stays finite.
        st = st.copy()
        same = st[:, 1] - st[:, 0] <= 0
        st[same, 1] = st[same, 0] + 1.0
        st[st[:, 3] <= 0, 3] = 1.0
        stats.append(st)
    W.write_scaling(os.path.join(d, "scaling.data"), stats)


def water_case(tmpdir, nmol=64, L=12.42, seed=1234, hidden=(25, 25),
               act="t t l", cutoff_type=1, rc=6.35, wseed=7):
    d = os.path.join(str(tmpdir), "water_%d_%d_%d" % (nmol, seed, cutoff_type))
    r, elem, box = W.water(nmol, L, seed)
    if not os.path.exists(os.path.join(d, "scaling.data")):
        W.write_model(d, W.water_settings(rc=rc, cutoff_type=cutoff_type,
                                          hidden=hidden, act=act),
                      W.WATER_ELEMENTS, (1, 8), hidden, seed=wseed)
        scale_from_oracle(d, r, elem, box)
    return d, r, elem, box


def mixed_case(tmpdir, natoms=64, seed=3):
    d = os.path.join(str(tmpdir), "mixed_%d_%d" % (natoms, seed))
    r, elem, box = W.random_structure(natoms, seed=seed)
    if not os.path.exists(os.path.join(d, "scaling.data")):
        W.write_model(d, W.mixed_settings(), W.MIXED_ELEMENTS, (5, 6, 7),
                      (30, 30), seed=11)
        scale_from_oracle(d, r, elem, box)
    return d, r, elem, box


def weighted_case(tmpdir, natoms=64, seed=5):
    """B/C/N structure with the element-weighted SF types 12 and 13 beside
    types 2 and 3."""
    d = os.path.join(str(tmpdir), "weighted_%d_%d" % (natoms, seed))
    r, elem, box = W.random_structure(natoms, seed=seed)
    if not os.path.exists(os.path.join(d, "scaling.data")):
        W.write_model(d, W.weighted_settings(), W.MIXED_ELEMENTS, (5, 6, 7),
                      (20, 20), seed=13)
        scale_from_oracle(d, r, elem, box)
    return d, r, elem, box


def compact_case(tmpdir, natoms=64, seed=7):
    """B/C/N structure with the compact SF families 20-25."""
    d = os.path.join(str(tmpdir), "compact_%d_%d" % (natoms, seed))
    r, elem, box = W.random_structure(natoms, seed=seed)
    if not os.path.exists(os.path.join(d, "scaling.data")):
        W.write_model(d, W.compact_settings(), W.MIXED_ELEMENTS, (5, 6, 7),
                      (20, 20), seed=17)
        scale_from_oracle(d, r, elem, box)
    return d, r, elem, box


def single_sf_model(tmpdir, name, elements, lines, cutoff_type=2,
                    cutoff_alpha=None, act="l l", all_sf=False):
    """Model with the given SF lines and a 1-neuron network; with the linear
    activations "l l" its output is the first SF (w = 1, b = 0):
    E_i = G_i[0]; with all_sf every SF enters with weight 1."""
    d = os.path.join(str(tmpdir), name)
    os.makedirs(d, exist_ok=True)
    syms = elements.split()
    ct = "cutoff_type %d" % cutoff_type
    if cutoff_alpha is not None:
        ct += " %.17g" % cutoff_alpha
    txt = ["number_of_elements %d" % len(syms), "elements " + elements, ct,
           "global_hidden_layers_short 1", "global_nodes_short 1",
           "global_activation_short " + act]
    txt += ["symfunction_short " + l for l in lines]
    with open(os.path.join(d, "input.nn"), "w") as f:
        f.write("\n".join(txt) + "\n")
    for s in syms:
        nsf = sum(1 for l in lines if l.split()[0] == s)
        z = W_atomic_number(s)
        w1 = [1.0] * nsf if all_sf else [1.0] + [0.0] * (nsf - 1)
        w = w1 + [0.0, 1.0, 0.0]
        with open(os.path.join(d, "weights.%03d.data" % z), "w") as f:
            f.write("".join("%.17g\n" % v for v in w))
    return d


_Z = {"H": 1, "He": 2, "B": 5, "C": 6, "N": 7, "O": 8, "S": 16, "Cu": 29,
      "U": 92}


def W_atomic_number(s):
    return _Z[s]
