"""Extracts the reference's own golden data into committed fixtures.

Run in a container that has the n2p2 sources at /root/reference:
    python tests/golden/make_golden.py
Writes (all data, no code):
  tests/golden/examples/<name>/{input.nn,scaling.data,weights.*.data,input.data}
      copies of examples/nnp-predict/<name> for the 2G examples whose SF and
      cutoff types this engine covers (H2O_RPBE-D3: G2/G3 + CT_TANHU;
      Cu2S_PBE: G2/G3/G9 + CT_POLY2, softplus, sigma scaling;
      Anisole_SCAN, DMABN_SCAN, Ethylbenzene_SCAN: the compact types 20 and
      21/22 with groups of 54 members);
  tests/golden/example_prediction.json
      energies and all force components of those examples from
      test/cpp/ExamplePrediction.h (tested there at 1000 eps);
  tests/golden/cutoff_functions.json
      f and df of all 9 cutoff types at r = 2.4, rc = 3.0, alpha = 0.2 from
      test/cpp/test_CutoffFunction.cpp (tested there at 100 eps);
  tests/golden/symfunc_fixtures.json
      the three-atom (one and two elements) and four-atom fixtures of
      test/cpp/fixtures.h with the setup lines and values of the SFs of
      atom 0 (types 2, 3, 9, 12, 13 and the compact types 20-25).
"""
import json
import os
import re
import shutil

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = ["H2O_RPBE-D3", "Cu2S_PBE", "Anisole_SCAN", "DMABN_SCAN",
            "Ethylbenzene_SCAN"]
NUM = r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?"


def examples():
    txt = open(os.path.join(REF, "test/cpp/ExamplePrediction.h")).read()
    res = {}
    for name in EXAMPLES:
        i = txt.index('ExamplePrediction("%s")' % name)
        j = txt.find("examples.push_back", i + 10)
        blk = txt[i:j if j > 0 else len(txt)]
        e = float(re.search(r"e->energy = (%s);" % NUM, blk).group(1))
        f = [[float(x) for x in m] for m in re.findall(
            r"Vec3D\(\s*(%s),\s*(%s),\s*(%s)\)" % (NUM, NUM, NUM), blk)]
        res[name] = {"energy": e, "forces": f}
        src = os.path.join(REF, "examples/nnp-predict", name)
        dst = os.path.join(OUT, "examples", name)
        os.makedirs(dst, exist_ok=True)
        for fn in sorted(os.listdir(src)):
            if fn in ("input.nn", "scaling.data", "input.data") or \
                    re.match(r"weights\.\d{3}\.data$", fn):
                shutil.copy(os.path.join(src, fn), os.path.join(dst, fn))
    with open(os.path.join(OUT, "example_prediction.json"), "w") as f:
        json.dump({"source": "test/cpp/ExamplePrediction.h",
                   "accuracy": "1000*eps absolute", "examples": res}, f,
                  indent=1)


def vec(txt, name):
    m = re.search(r"%s\s*=\s*\{(.*?)\};" % re.escape(name), txt, re.S)
    return m.group(1)


def cutoffs():
    txt = open(os.path.join(REF, "test/cpp/test_CutoffFunction.cpp")).read()
    f = [float(x) for x in re.findall(NUM, vec(txt, "vector<double> const f"))]
    df = [float(x) for x in re.findall(NUM, vec(txt, "vector<double> const df"))]
    with open(os.path.join(OUT, "cutoff_functions.json"), "w") as fh:
        json.dump({"source": "test/cpp/test_CutoffFunction.cpp",
                   "rc": 3.0, "alpha": 0.2, "r": 2.4,
                   "types": list(range(9)), "f": f, "df": df,
                   "accuracy": "100*eps absolute"}, fh, indent=1)


def fixtures():
    txt = open(os.path.join(REF, "test/cpp/fixtures.h")).read()
    out = {"source": "test/cpp/fixtures.h", "cutoff_type": 2, "rc": 10.0,
           "accuracy": "1000*eps absolute", "sets": {}}
    # Positions and elements of the fixtures (atom 0 at the origin).
    structs = {
        "Mono": {"elements": "H", "r": [[0, 0, 0], [1, 2, 1], [0, 3, -1]],
                 "elem": [0, 0, 0]},
        "Dual": {"elements": "H He", "r": [[0, 0, 0], [1, 2, 1], [0, 3, -1]],
                 "elem": [0, 1, 1]},
        "Mixed": {"elements": "S Cu U",
                  "r": [[0, 0, 0], [1, 2, 1], [0, 3, -1], [-2, 1, -0.5]],
                  "elem": [0, 0, 1, 2]},
    }
    for key, st in structs.items():
        types = [int(x) for x in re.findall(r"\d+", vec(txt, "vector<size_t> const types" + key))]
        lines = re.findall(r'"([^"]*)"', vec(txt, "vector<string> const setupLines" + key))
        vals = [float(x) for x in re.findall(NUM, vec(txt, "vector<double> const values" + key))]
        keep = [(t, l, v) for t, l, v in zip(types, lines, vals) if t in (2, 3, 9, 12, 13, 20, 21, 22, 23, 24, 25)]
        st = dict(st)
        st["sf"] = [{"type": t, "line": l, "value": v} for t, l, v in keep]
        out["sets"][key] = st
    # Check the positions stated above against the fixture source text.
    assert "r[0] = -2.0" in txt and "r[2] = -0.5" in txt and "r[1] = 3.0" in txt
    with open(os.path.join(OUT, "symfunc_fixtures.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    examples()
    cutoffs()
    fixtures()
    print("golden fixtures written to", OUT)
