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

The SF histograms (Dataset::writeSymmetryFunctionHistograms,
Dataset.cpp:1103-1240) come from a second pass over the batches:
- the raw G of each batch are evaluated on the device;
- they are binned on the host into <nbins> uniform bins over
  [min, max + (max - min) / (nbins - 1)] (the reference's GSL histogram;
  a value lands in bin i when edge_i <= G < edge_(i+1));
- each element and SF gets a file sf.<Z>.<SF index>.histo, with the SF's
  #SFINFO parameter lines (SymFnc*::parameterInfo, 12-wide names), its
  statistics and the column header.

Usage, in a directory with input.nn and input.data (as nnp-scaling):

    python -m n2p2_amd.scaling <nbins>
"""
import math
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


def statistics(settings, structures, max_atoms=4_000_000, device=0,
               raw=False):
    """Per element [nsf x 4] of min, max, mean, sigma over all atoms of the
    structures [(r, elem, box or None), ...], in the model's internal units
    (positions converted with conv_length when the settings normalise, as
    nnp-scaling's toNormalizedUnits).  raw: also return the per-element
    atom counts and [nsf x 4] accumulators (min, max, sum, sum of
    squares)."""
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
    if raw:
        return out, cnt, acc
    return out


def _pad(name, width=12):
    """utility.cpp:79-88 pad (right-filled with blanks)."""
    return name + " " * max(0, width - len(name))


def sf_info(model, settings_lines, e, symbols):
    """#SFINFO lines of every SF of element e in SF order: the
    parameterInfo of SymFnc.cpp:32-50, SymFncBaseCutoff.cpp:24-36,
    SymFncExpRad.cpp:201-215, SymFncBaseExpAng.cpp:143-166, the weighted
    SymFncExpRadWeighted / SymFncExpAngnWeighted, SymFncBaseComp.cpp:25-37,
    SymFncCompRad.cpp:198-207 and SymFncBaseCompAng(Weighted).  Lengths are
    printed as the reference does, converted back from internal units."""
    _, _, _, conv = model.normalization()
    if not model.normalization()[0]:
        conv = 1.0
    table = model.sf_table(e)
    out = []
    for idx, row in enumerate(table):
        t, e1, e2, eta, rs, lam, zeta, rc, ct, alpha, _, ln = row
        t, ln = int(t), int(ln)
        w = settings_lines[ln].split("#", 1)[0].split()[1:]   # after keyword
        v = [(_pad("lineNumber") + "%d") % (ln + 1),
             (_pad("index") + "%d") % (idx + 1),
             (_pad("type") + "%d") % t,
             (_pad("ec") + "%s") % symbols[e],
             (_pad("rc") + "%14.8E") % (rc / conv)]
        if t in (2, 3, 9, 12, 13):
            v += [(_pad("subtype") + "ct%d") % int(ct),
                  (_pad("alpha") + "%14.8E") % alpha]
            if t == 2:
                v.append((_pad("e1") + "%s") % symbols[int(e1)])
            if t in (3, 9):
                v += [(_pad("e1") + "%s") % symbols[int(e1)],
                      (_pad("e2") + "%s") % symbols[int(e2)]]
            v.append((_pad("eta") + "%14.8E") % (eta * conv * conv))
            if t in (3, 9, 13):
                v += [(_pad("lambda") + "%14.8E") % lam,
                      (_pad("zeta") + "%14.8E") % zeta]
            v.append((_pad("rs") + "%14.8E") % (rs / conv))
        else:
            # compact: subtype and rl from the settings line
            p = 2
            if t in (20, 21, 22):
                p += 1 if t == 20 else 2
            rl = float(w[p]) * conv
            ang = (w[p + 2], w[p + 3]) if t not in (20, 23) else None
            sub = w[p + 4] if ang else w[p + 2]
            v += [(_pad("subtype") + "%s") % sub,
                  (_pad("rl") + "%14.8E") % (rl / conv)]
            if t == 20:
                v.append((_pad("e1") + "%s") % symbols[int(e1)])
            if t in (21, 22):
                v += [(_pad("e1") + "%s") % symbols[int(e1)],
                      (_pad("e2") + "%s") % symbols[int(e2)]]
            if ang:
                v += [(_pad("angleLeft") + "%14.8E") % float(ang[0]),
                      (_pad("angleRight") + "%14.8E") % float(ang[1])]
        out.append(v)
    return out


HISTO_COLUMNS = (
    [16, 16, 16],
    ["symfunc_l", "symfunc_r", "hist"],
    ["Symmetry function value, left bin limit.",
     "Symmetry function value, right bin limit.", "Histogram count."])


def histogram_edges(lo, hi, nbins):
    """The reference's bins: nbins uniform bins over
    [lo, hi + (hi - lo) / (nbins - 1)] (Dataset.cpp:1113-1127), edges as a
    uniform GSL histogram sets them, lo + (i / n) (top - lo)."""
    nb = nbins - 1
    top = hi + (hi - lo) / nb
    n = nb + 1
    return np.array([lo + (i / n) * (top - lo) for i in range(n + 1)])


def write_histograms(settings, structures, nbins, directory=".",
                     max_atoms=4_000_000, device=0,
                     fmt="sf.%03d.%04d.histo"):
    """SF histogram files of nnp-scaling for the structures; returns the
    number of files written (SFs whose minimum equals their maximum get
    none, as in the reference)."""
    import os
    model = engine.Model.load(settings, None, "")
    used, _, _, conv_length = model.normalization()
    if not used:
        conv_length = 1.0
    stats, cnt, acc = statistics(settings, structures, max_atoms, device,
                                 raw=True)
    ne = model.num_elements
    edges = [[None] * model.num_sf(k) for k in range(ne)]
    counts = [np.zeros((model.num_sf(k), nbins)) for k in range(ne)]
    for k in range(ne):
        if acc[k] is None:
            continue
        for s_ in range(model.num_sf(k)):
            lo, hi = acc[k][s_, 0], acc[k][s_, 1]
            if lo < hi:
                edges[k][s_] = histogram_edges(lo, hi, nbins)
    ctx = engine.Context(model, device=device)
    for batch in _batches(structures, max_atoms):
        conv = []
        for r, e, box in batch:
            if conv_length != 1.0:
                r, box = nio.to_normalized(r, box, conv_length)
            conv.append((r, e, box))
        ctx.set_structures(conv)
        ctx.calculate_neighbor_list(model.max_cutoff)
        ctx.calculate_symmetry_function_groups(derivatives=False)
        G = ctx.G()
        el = np.concatenate([np.asarray(e) for _, e, _ in conv])
        for k in range(ne):
            gk = G[el == k][:, :model.num_sf(k)]
            for s_ in range(model.num_sf(k)):
                ed = edges[k][s_]
                if ed is None:
                    continue
                i = np.searchsorted(ed, gk[:, s_], side="right") - 1
                i = i[(i >= 0) & (i < nbins)]
                counts[k][s_] += np.bincount(i, minlength=nbins)
    ctx.close()
    lines = open(settings).read().split("\n")
    symbols = [None] * ne
    for l in lines:
        w = l.split("#", 1)[0].split()
        if w and w[0] == "elements":
            symbols = w[1:1 + ne]
    size, name, info = HISTO_COLUMNS
    nfiles = 0
    for k in range(ne):
        infos = sf_info(model, lines, k, symbols)
        z = model.atomic_number(k)
        for s_ in range(model.num_sf(k)):
            ed = edges[k][s_]
            if ed is None:
                continue
            n = float(cnt[k])
            a = acc[k][s_]
            path = os.path.join(directory, fmt % (z, s_ + 1))
            with open(path, "w") as f:
                for l in infos[s_]:
                    f.write("#SFINFO %s\n" % l)
                f.write("#SFINFO min         %15.8E\n" % a[0])
                f.write("#SFINFO max         %15.8E\n" % a[1])
                f.write("#SFINFO mean        %15.8E\n" % (a[2] / n))
                f.write("#SFINFO sigma       %15.8E\n"
                        % math.sqrt(1.0 / (n - 1) * (a[3] - a[2] * a[2] / n)))
                f.writelines(file_header(["Symmetry function histogram."],
                                         size, name, info))
                for i in range(nbins):
                    f.write("%16.8E %16.8E %16.8E\n"
                            % (ed[i], ed[i + 1], counts[k][s_, i]))
            nfiles += 1
    return nfiles


def main(argv=None):
    argv = sys.argv if argv is None else argv
    if len(argv) != 2:
        print("USAGE: %s <nbins>\n"
              "       <nbins> ... Number of symmetry function histogram bins.\n"
              "       Execute in directory with these NNP files present:\n"
              "       - input.data (structure file)\n"
              "       - input.nn (NNP settings)" % argv[0])
        return 1
    model = engine.Model.load("input.nn", None, "")
    structures = [(r, e, b) for r, e, b, _ in
                  nio.read_structures("input.data", model.element_index)]
    stats = statistics("input.nn", structures)
    write_scaling_file("scaling.data", stats)
    nfiles = write_histograms("input.nn", structures, int(argv[1]))
    print("Wrote scaling.data: %d structures, %s SFs per element; %d "
          "histogram files." % (len(structures), [len(s) for s in stats],
                                nfiles))
    return 0


if __name__ == "__main__":
    sys.exit(main())
