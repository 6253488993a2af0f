"""nnp-scaling on the GPU: symmetry-function statistics of a data set and
the scaling.data file (nnp-scaling.cpp, Dataset.cpp:974-1100).

The reference evaluates the structures of input.data one at a time
(neighbour list, symmetry functions without scaling) and accumulates per
element and SF the minimum, maximum, sum and sum of squares
(SymFncStatistics, SymFncStatistics.h:28-147); mean = sum / n and
sigma = sqrt((sum2 - sum^2 / n) / (n - 1)) are written with a fixed header.
Here the structures are loaded as batches (hdnnp_set_structures), one
batch evaluated per launch, and the accumulators of the batches are
combined on the host.

Usage, in a directory with input.nn and input.data (as nnp-scaling):

    python -m n2p2_amd.scaling <nbins>

<nbins> is accepted for the command line of nnp-scaling; the SF histogram
files (Dataset::writeSymmetryFunctionHistograms) are not written.
"""
import sys

import numpy as np

from n2p2_amd import engine
from n2p2_amd import io as nio


def file_header(title, col_size, col_name, col_info, comment="#"):
    """The comment header of n2p2 data files (utility.cpp:111-221
    createFileHeader): title block, column table, column numbers and
    names right-aligned to the column widths."""
    std = 80
    sep = comment * std + "\n"
    h = [sep] + ["%s %s\n" % (comment, t) for t in title] + [sep]
    wn = max(len(n) for n in col_name)
    h.append("%s %-4s %-*s Description\n" % (comment, "Col", wn, "Name"))
    h.append(sep)
    for i, (n, info) in enumerate(zip(col_name, col_info)):
        h.append("%s %-4d %-*s %s\n" % (comment, i + 1, wn, n, info))
    wd = sum(s + 1 for s in col_size) - 1
    h.append(comment * wd + "\n" if wd > std else sep)
    nums, names = comment, comment
    for i, (s, n) in enumerate(zip(col_size, col_name)):
        w = s - 2 if i == 0 else s
        nums += " %*d" % (w, i + 1)
        names += " %*s" % (w, n[:w])
    h += [nums + "\n", names + "\n", comment * wd + "\n"]
    return h


SCALING_COLUMNS = (
    [10, 10, 24, 24, 24, 24],
    ["e_index", "sf_index", "sf_min", "sf_max", "sf_mean", "sf_sigma"],
    ["Element index.", "Symmetry function index.",
     "Symmetry function minimum.", "Symmetry function maximum.",
     "Symmetry function mean.", "Symmetry function sigma."])


def write_scaling_file(path, stats):
    """scaling.data as Dataset::writeSymmetryFunctionScaling writes it
    (Dataset.cpp:1004-1100).  stats: per element an array [nsf x 4] of
    min, max, mean, sigma."""
    size, name, info = SCALING_COLUMNS
    with open(path, "w") as f:
        f.writelines(file_header(["Symmetry function scaling data."], size,
                                 name, info))
        for e, st in enumerate(stats):
            for i, (mn, mx, mean, sig) in enumerate(st):
                f.write("%10d %10d %24.16E %24.16E %24.16E %24.16E\n"
                        % (e + 1, i + 1, mn, mx, mean, sig))


def _batches(structures, max_atoms):
    cur, n = [], 0
    for s in structures:
        if cur and n + len(s[1]) > max_atoms:
            yield cur
            cur, n = [], 0
        cur.append(s)
        n += len(s[1])
    if cur:
        yield cur


def statistics(settings, structures, max_atoms=4_000_000, device=0):
    """Per element [nsf x 4] of min, max, mean, sigma over all atoms of the
    structures [(r, elem, box or None), ...], in the model's internal units
    (positions converted with conv_length when the settings normalise, as
    nnp-scaling's toNormalizedUnits)."""
    model = engine.Model.load(settings, None, "")
    used, _, _, conv_length = model.normalization()
    if not used:
        conv_length = 1.0
    ctx = engine.Context(model, device=device)
    ne = model.num_elements
    acc = [None] * ne
    cnt = [0] * ne
    for batch in _batches(structures, max_atoms):
        conv = []
        for r, e, box in batch:
            if conv_length != 1.0:
                r, box = nio.to_normalized(r, box, conv_length)
            conv.append((r, e, box))
        ctx.set_structures(conv)
        ctx.calculate_neighbor_list(model.max_cutoff)
        for k in range(ne):
            n, st = ctx.sf_statistics_raw(k)
            if n == 0:
                continue
            if acc[k] is None:
                acc[k] = st.copy()
            else:
                acc[k][:, 0] = np.minimum(acc[k][:, 0], st[:, 0])
                acc[k][:, 1] = np.maximum(acc[k][:, 1], st[:, 1])
                acc[k][:, 2] += st[:, 2]
                acc[k][:, 3] += st[:, 3]
            cnt[k] += n
    ctx.close()
    out = []
    for k in range(ne):
        nsf = model.num_sf(k)
        if acc[k] is None:
            out.append(np.zeros((nsf, 4)))
            continue
        n = float(cnt[k])
        a = acc[k]
        mean = a[:, 2] / n
        sigma = np.sqrt((a[:, 3] - a[:, 2] * a[:, 2] / n) / (n - 1.0))
        out.append(np.stack([a[:, 0], a[:, 1], mean, sigma], axis=1))
    return out


def main(argv=None):
    argv = sys.argv if argv is None else argv
    if len(argv) != 2:
        print("USAGE: %s <nbins>\n"
              "       <nbins> ... Number of symmetry function histogram bins "
              "(histograms are not written).\n"
              "       Execute in directory with these NNP files present:\n"
              "       - input.data (structure file)\n"
              "       - input.nn (NNP settings)" % argv[0])
        return 1
    model = engine.Model.load("input.nn", None, "")
    structures = [(r, e, b) for r, e, b, _ in
                  nio.read_structures("input.data", model.element_index)]
    stats = statistics("input.nn", structures)
    write_scaling_file("scaling.data", stats)
    print("Wrote scaling.data: %d structures, %s SFs per element."
          % (len(structures), [len(s) for s in stats]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
