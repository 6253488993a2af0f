"""Synthetic workloads: structures, settings files and random-init weights.

Everything is generated from a seed with numpy, written in n2p2's file
formats (input.nn keywords, input.data, weights.%03zu.data in
NeuralNetwork::setConnections order, scaling.data in the layout of
Dataset::writeSymmetryFunctionScaling).

Workloads
  water(nmol, L): water-like H/O box; O on a ceil(cbrt(nmol))^3 lattice
      (nmol sites picked by a seeded permutation when the lattice has more),
      jittered by U(+-0.3 A); each O carries two H at 0.96 A with a 104.5 deg
      HOH angle and a uniformly random orientation.
        192 atoms:     nmol=64,    L=12.42  (4x4x4 lattice)
        24,576 atoms:  nmol=8192,  L=62.1
        196,608 atoms: nmol=65536, L=124.2
  water_settings(): 8 radial G2 per neighbour element (eta log-spaced
      0.001-0.5 A^-2, rs=0) and 8 angular G3 per element pair (lambda=+-1,
      zeta in {1,2,4,16}, eta=0.001), rc=6.35 A, cosine cutoff, min/max
      scaling to [0,1], network nsf-25-25-1 with tanh hidden layers.
  random_structures(): uneven three-element (B/C/N) periodic boxes with
      random positions at 0.08 atoms/A^3 and minimum pair distance 1.8 A.
  mixed_settings(): 10 radial G2 per element + 12 wide G9 per element pair,
      rc=6.0 A, polynomial cutoff (cutoff_type 6), softplus network 30-30.
"""

import math
import os

import numpy as np

WATER_ELEMENTS = ("H", "O")          # element index 0 = H, 1 = O (sorted by Z)
MIXED_ELEMENTS = ("B", "C", "N")


def _rotation(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def remap(r, box):
    """Structure::remap (Structure.cpp:1223-1247) for an orthorhombic box
    given as 3x3 rows: wrap into [0, 1) fractional coordinates."""
    box = np.asarray(box, dtype=np.float64)
    inv = np.linalg.inv(box)
    f = r @ inv
    f = f - np.floor(f)
    f[f >= 1.0] = 0.0
    return f @ box


def water(nmol, L, seed=1234):
    """Positions (N x 3), element indices (N, H=0, O=1) and box (3 x 3)."""
    rng = np.random.default_rng(seed)
    n = int(math.ceil(round(nmol ** (1.0 / 3.0), 9)))
    a = L / n
    g = np.arange(n)
    sites = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    sites = sites.astype(np.float64) * a + a / 2
    if len(sites) > nmol:
        sites = sites[np.sort(rng.permutation(len(sites))[:nmol])]
    O = sites + rng.uniform(-0.3, 0.3, sites.shape)
    th = math.radians(104.5)
    q = rng.normal(size=(nmol, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    h1 = np.array([0.96, 0.0, 0.0])
    h2 = np.array([0.96 * math.cos(th), 0.96 * math.sin(th), 0.0])
    R = _rotation(q.T)                       # 3 x 3 x nmol
    r = np.empty((nmol, 3, 3))
    r[:, 0] = O
    r[:, 1] = O + np.einsum("abm,b->ma", R, h1)
    r[:, 2] = O + np.einsum("abm,b->ma", R, h2)
    r = r.reshape(-1, 3)
    elem = np.tile(np.array([1, 0, 0], dtype=np.int32), nmol)
    box = np.diag([L, L, L]).astype(np.float64)
    return remap(r, box), elem, box


def water_settings(rc=6.35, cutoff_type=1, scale=True, hidden=(25, 25),
                   act="t t l"):
    el = WATER_ELEMENTS
    lines = ["number_of_elements 2", "elements H O",
             "cutoff_type %d" % cutoff_type]
    if scale:
        lines += ["scale_symmetry_functions", "scale_min_short 0.0",
                  "scale_max_short 1.0"]
    lines += ["global_hidden_layers_short %d" % len(hidden),
              "global_nodes_short " + " ".join(str(h) for h in hidden),
              "global_activation_short " + act]
    for e in el:
        for e1 in el:
            for eta in np.geomspace(0.001, 0.5, 8):
                lines.append("symfunction_short %s 2 %s %.10e 0.0 %.4f"
                             % (e, e1, eta, rc))
        for i1 in range(2):
            for i2 in range(i1, 2):
                for lam in (1.0, -1.0):
                    for zeta in (1.0, 2.0, 4.0, 16.0):
                        lines.append("symfunction_short %s 3 %s %s 0.001 %.1f"
                                     " %.1f %.4f" % (e, el[i1], el[i2], lam,
                                                     zeta, rc))
    return "\n".join(lines) + "\n"


def mixed_settings(rc=6.0, cutoff_type=6):
    el = MIXED_ELEMENTS
    lines = ["number_of_elements 3", "elements B C N",
             "cutoff_type %d 0.0" % cutoff_type,
             "scale_symmetry_functions", "scale_min_short 0.0",
             "scale_max_short 1.0",
             "global_hidden_layers_short 2", "global_nodes_short 30 30",
             "global_activation_short p p l"]
    for e in el:
        for e1 in el:
            for eta in np.geomspace(0.002, 1.0, 10):
                lines.append("symfunction_short %s 2 %s %.10e 0.0 %.4f"
                             % (e, e1, eta, rc))
        for i1 in range(3):
            for i2 in range(i1, 3):
                for eta in (0.001, 0.01, 0.05):
                    for lam in (1.0, -1.0):
                        for zeta in (1.0, 4.0):
                            lines.append("symfunction_short %s 9 %s %s %.4f "
                                         "%.1f %.1f %.4f" % (e, el[i1], el[i2],
                                                             eta, lam, zeta, rc))
    return "\n".join(lines) + "\n"


def weighted_settings(rc=6.0, cutoff_type=6):
    """B/C/N settings that mix the element-resolved types (2, 3) with the
    element-weighted types 12 and 13 (SymFncExpRadWeighted,
    SymFncExpAngnWeighted): two cutoff radii, shifted and unshifted
    Gaussians, integer and non-integer zeta."""
    el = MIXED_ELEMENTS
    lines = ["number_of_elements 3", "elements B C N",
             "cutoff_type %d 0.0" % cutoff_type,
             "scale_symmetry_functions", "scale_min_short 0.0",
             "scale_max_short 1.0",
             "global_hidden_layers_short 2", "global_nodes_short 20 20",
             "global_activation_short t t l"]
    for e in el:
        for e1 in el:
            for eta in (0.01, 0.1, 0.5):
                lines.append("symfunction_short %s 2 %s %.4f 0.0 %.4f"
                             % (e, e1, eta, rc))
        for e1, e2 in ((el[0], el[1]), (el[1], el[2])):
            for lam in (1.0, -1.0):
                for zeta in (1.0, 4.0):
                    lines.append("symfunction_short %s 3 %s %s 0.01 %.1f %.1f "
                                 "%.4f" % (e, e1, e2, lam, zeta, rc))
        for rcw in (4.5, rc):
            for eta, rs in ((0.02, 0.0), (0.3, 0.0), (0.5, 1.5), (1.0, 2.5)):
                lines.append("symfunction_short %s 12 %.4f %.4f %.4f"
                             % (e, eta, rs, rcw))
        for eta, rs in ((0.005, 0.0), (0.05, 1.2)):
            for lam in (1.0, -1.0):
                for zeta in (1.0, 2.0, 3.5, 16.0):
                    lines.append("symfunction_short %s 13 %.4f %.4f %.1f %.1f "
                                 "%.4f" % (e, eta, rs, lam, zeta, rc))
    return "\n".join(lines) + "\n"


def compact_settings(rc=6.0):
    """B/C/N settings with the compact (polynomial) families 20-25 beside a
    few G2: every core (p1-p4, asymmetric p2a, e), negative rl, angle
    intervals centred on 0 and 180 degrees, and within each angular group
    two members with equal (rl, rc) and different cores next to each other
    in sort order."""
    el = MIXED_ELEMENTS
    lines = ["number_of_elements 3", "elements B C N",
             "cutoff_type 6 0.0",
             "scale_symmetry_functions", "scale_min_short 0.0",
             "scale_max_short 1.0",
             "global_hidden_layers_short 2", "global_nodes_short 20 20",
             "global_activation_short t t l"]
    rad = ((-5.0, 5.0, "p2"), (0.0, 6.0, "p3"), (1.5, 6.0, "p2a"),
           (0.5, 5.5, "e"))
    ang = ((0.0, 4.5, 0.0, 180.0, "p2"), (0.0, 4.5, 30.0, 120.0, "p3"),
           (1.0, 6.0, 60.0, 180.0, "p3"), (-6.0, 6.0, -90.0, 90.0, "p1"),
           (0.0, 5.0, 90.0, 270.0, "p4"))
    for e in el:
        for e1 in el:
            lines.append("symfunction_short %s 2 %s 0.1 0.0 %.4f" % (e, e1, rc))
            for rl, rr, st in rad:
                lines.append("symfunction_short %s 20 %s %.2f %.2f %s"
                             % (e, e1, rl, rr, st))
        for e1, e2 in ((el[0], el[1]), (el[1], el[2]), (el[2], el[2])):
            for t in (21, 22):
                for rl, rr, al, ar, st in ang:
                    lines.append("symfunction_short %s %d %s %s %.2f %.2f "
                                 "%.1f %.1f %s" % (e, t, e1, e2, rl, rr, al,
                                                   ar, st))
        for rl, rr, st in ((-5.0, 5.0, "p2"), (0.0, 6.0, "e"), (1.0, 6.0, "p4")):
            lines.append("symfunction_short %s 23 %.2f %.2f %s"
                         % (e, rl, rr, st))
        for t in (24, 25):
            for rl, rr, al, ar, st in ((0.0, 4.5, 0.0, 180.0, "p2"),
                                       (0.5, 6.0, 45.0, 135.0, "p3"),
                                       (-6.0, 6.0, -60.0, 60.0, "p2a")):
                lines.append("symfunction_short %s %d %.2f %.2f %.1f %.1f %s"
                             % (e, t, rl, rr, al, ar, st))
    return "\n".join(lines) + "\n"


def count_sf(settings_text, symbol):
    return sum(1 for l in settings_text.splitlines()
               if l.startswith("symfunction_short %s " % symbol))


def xavier_weights(sizes, seed):
    """Weights in NeuralNetwork::setConnections order (NeuralNetwork.cpp:
    147-169): per layer, for each input j, for each neuron k: W[k][j]; then
    the layer's biases.  W ~ U(+-sqrt(6 / (fan_in + fan_out))), b = 0."""
    rng = np.random.default_rng(seed)
    out = []
    for a, b in zip(sizes[:-1], sizes[1:]):
        lim = math.sqrt(6.0 / (a + b))
        w = rng.uniform(-lim, lim, (b, a))     # W[k][j]
        out.extend(w.T.reshape(-1))           # j outer, k inner
        out.extend([0.0] * b)
    return np.array(out)


def write_model(d, settings_text, elements, z_numbers, hidden, seed=7):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "input.nn"), "w") as f:
        f.write(settings_text)
    for k, (sym, z) in enumerate(zip(elements, z_numbers)):
        nsf = count_sf(settings_text, sym)
        w = xavier_weights([nsf] + list(hidden) + [1], seed + 101 * k)
        with open(os.path.join(d, "weights.%03d.data" % z), "w") as f:
            f.write("".join("%.16e\n" % v for v in w))


def write_scaling(path, stats_per_element):
    """scaling.data (Dataset.cpp:1074): e_index sf_index min max mean sigma,
    1-based indices, elements in element-map order."""
    with open(path, "w") as f:
        f.write("#" + "#" * 79 + "\n# Symmetry function scaling data.\n#"
                + "#" * 79 + "\n")
        for e, st in enumerate(stats_per_element):
            for i, (mn, mx, mean, sig) in enumerate(st):
                f.write("%10d %10d %24.16E %24.16E %24.16E %24.16E\n"
                        % (e + 1, i + 1, mn, mx, mean, sig))


def write_structure(path, r, elem, box, symbols):
    with open(path, "w") as f:
        f.write("begin\n")
        if box is not None:
            for v in box:
                f.write("lattice %.16e %.16e %.16e\n" % tuple(v))
        for p, e in zip(r, elem):
            f.write("atom %.16e %.16e %.16e %s 0.0 0.0 0.0 0.0 0.0\n"
                    % (p[0], p[1], p[2], symbols[e]))
        f.write("energy 0.0\ncharge 0.0\nend\n")


def random_structure(natoms, density=0.08, dmin=1.8, fractions=(0.4, 0.4, 0.2),
                     seed=0, max_tries=200000):
    """Random sequential addition in a cubic periodic box; element labels by
    the given fractions.  Returns r, elem, box."""
    rng = np.random.default_rng(seed)
    L = (natoms / density) ** (1.0 / 3.0)
    pts = np.empty((0, 3))
    tries = 0
    while len(pts) < natoms:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random_structure: could not place atoms")
        p = rng.uniform(0.0, L, 3)
        if len(pts):
            d = pts - p
            d -= L * np.round(d / L)
            if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                continue
        pts = np.vstack([pts, p])
    counts = np.floor(np.array(fractions) * natoms).astype(int)
    counts[0] += natoms - counts.sum()
    elem = np.concatenate([np.full(c, k, dtype=np.int32)
                           for k, c in enumerate(counts)])
    rng.shuffle(elem)
    box = np.diag([L, L, L]).astype(np.float64)
    return remap(pts, box), elem, box


def batch_structures(nstruct=512, lo=16, hi=512, seed=2024):
    """Batch of independent three-element periodic structures with atom
    counts drawn uniformly from [lo, hi] (random_structure each: density
    0.08 atoms/A^3, fractions 0.4/0.4/0.2).  Returns [(r, elem, box), ...]."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, nstruct)
    return [random_structure(int(n), seed=seed * 1000 + k)
            for k, n in enumerate(sizes)]


def sheared(r, box, shear):
    """The structure under the linear map x -> x @ shear (rows of box are
    lattice vectors): fractional coordinates, hence the remap into the
    cell, are unchanged."""
    S = np.asarray(shear, dtype=np.float64)
    return r @ S, box @ S


KB_EV = 8.617333262e-5               # Boltzmann constant, eV/K
FTOA = 9.648533212e-3                # (eV/A)/amu in A/fs^2
WATER_MASSES = (1.008, 15.999)       # amu, element order of WATER_ELEMENTS


def maxwell_boltzmann(masses, T, seed=0):
    """Velocities (A/fs) at temperature T for atoms of the given masses
    (amu): normal components with variance kB T / m, total momentum
    removed."""
    m = np.asarray(masses, dtype=np.float64)
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(len(m), 3)) * np.sqrt(KB_EV * T * FTOA / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    return v


WATER_CONFIGS = {
    "water192": (64, 12.42),
    "water24576": (8192, 62.1),
    "water196608": (65536, 124.2),
}
