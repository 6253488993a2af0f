"""n2p2 data-file I/O: input.data structures and unit conversion.

read_structure follows Structure::readFromLines (Structure.cpp:142-271):
the first configuration between "begin" and "end", "lattice" rows as box
vectors, "atom x y z element charge _ fx fy fz" lines; periodic positions are
wrapped with Structure::remap (Structure.cpp:1223-1247) using the inverse box
of Structure::calculateInverseBox (Structure.cpp:519-545), evaluated in the
same operation order so positions match the reference bit for bit.
"""

import numpy as np


def inverse_box(b):
    b = np.asarray(b, dtype=np.float64)
    invdet = (b[0, 0] * b[1, 1] * b[2, 2] + b[1, 0] * b[2, 1] * b[0, 2]
              + b[2, 0] * b[0, 1] * b[1, 2] - b[2, 0] * b[1, 1] * b[0, 2]
              - b[0, 0] * b[2, 1] * b[1, 2] - b[1, 0] * b[0, 1] * b[2, 2])
    invdet = 1.0 / invdet
    ib = np.empty((3, 3))
    ib[0, 0] = b[1, 1] * b[2, 2] - b[2, 1] * b[1, 2]
    ib[0, 1] = b[2, 1] * b[0, 2] - b[0, 1] * b[2, 2]
    ib[0, 2] = b[0, 1] * b[1, 2] - b[1, 1] * b[0, 2]
    ib[1, 0] = b[2, 0] * b[1, 2] - b[1, 0] * b[2, 2]
    ib[1, 1] = b[0, 0] * b[2, 2] - b[2, 0] * b[0, 2]
    ib[1, 2] = b[1, 0] * b[0, 2] - b[0, 0] * b[1, 2]
    ib[2, 0] = b[1, 0] * b[2, 1] - b[2, 0] * b[1, 1]
    ib[2, 1] = b[2, 0] * b[0, 1] - b[0, 0] * b[2, 1]
    ib[2, 2] = b[0, 0] * b[1, 1] - b[1, 0] * b[0, 1]
    return ib * invdet


def remap(r, box):
    """Structure::remap for every atom (vectorised, same operation order)."""
    r = np.asarray(r, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64)
    ib = inverse_box(box)
    f = r[:, 0:1] * ib[0] + r[:, 1:2] * ib[1]
    f = f + r[:, 2:3] * ib[2]
    trunc = np.trunc(f)
    f = np.where(f > 1.0, f - trunc, f)
    trunc = np.trunc(f)
    f = np.where(f < 0.0, f + (1.0 - trunc), f)
    f = np.where(f == 1.0, 0.0, f)
    out = f[:, 0:1] * box[0] + f[:, 1:2] * box[1]
    return out + f[:, 2:3] * box[2]


def read_structure(path, element_index):
    """Returns (r, elem, box or None, energy_ref).  element_index maps a
    chemical symbol to the model's element index (ElementMap)."""
    r, elem, box = [], [], []
    energy = 0.0
    begun = False
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w:
                continue
            if w[0] == "begin":
                begun = True
                continue
            if not begun or w[0] == "comment":
                continue
            if w[0] == "lattice":
                box.append([float(x) for x in w[1:4]])
            elif w[0] == "atom":
                r.append([float(x) for x in w[1:4]])
                e = element_index(w[4])
                if e < 0:
                    raise RuntimeError("ERROR: Unknown element \"%s\".\n" % w[4])
                elem.append(e)
            elif w[0] == "energy":
                energy = float(w[1])
            elif w[0] == "charge":
                pass
            elif w[0] == "end":
                break
            else:
                raise RuntimeError("ERROR: Unexpected file content, unknown "
                                   "keyword \"%s\".\n" % w[0])
    if len(box) not in (0, 3):
        raise RuntimeError("ERROR: Strange number of box vectors.\n")
    r = np.array(r, dtype=np.float64).reshape(-1, 3)
    elem = np.array(elem, dtype=np.int32)
    b = np.array(box, dtype=np.float64) if box else None
    if b is not None:
        r = remap(r, b)
    return r, elem, b, energy


def read_structures(path, element_index, with_forces=False):
    """Every structure of an input.data file (Structure::readFromFile over
    the whole file, Dataset::readFromFile): a list of (r, elem, box or None,
    energy_ref), positions remapped into the cell; with_forces appends the
    reference forces (atom columns 7-9, Structure.cpp:214-232) to each
    tuple."""
    out, block = [], []
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w:
                continue
            block.append(line)
            if w[0] == "end":
                s = _parse_block(block, element_index)
                out.append(s if with_forces else s[:4])
                block = []
    return out


def _parse_block(lines, element_index):
    r, elem, box, force = [], [], [], []
    energy = 0.0
    for line in lines:
        w = line.split()
        if w[0] in ("begin", "comment", "charge", "end"):
            continue
        if w[0] == "lattice":
            box.append([float(x) for x in w[1:4]])
        elif w[0] == "atom":
            r.append([float(x) for x in w[1:4]])
            e = element_index(w[4])
            if e < 0:
                raise RuntimeError("ERROR: Unknown element \"%s\".\n" % w[4])
            elem.append(e)
            force.append([float(x) for x in w[7:10]] if len(w) >= 10
                         else [0.0, 0.0, 0.0])
        elif w[0] == "energy":
            energy = float(w[1])
        else:
            raise RuntimeError("ERROR: Unexpected file content, unknown "
                               "keyword \"%s\".\n" % w[0])
    if len(box) not in (0, 3):
        raise RuntimeError("ERROR: Strange number of box vectors.\n")
    r = np.array(r, dtype=np.float64).reshape(-1, 3)
    elem = np.array(elem, dtype=np.int32)
    b = np.array(box, dtype=np.float64) if box else None
    if b is not None:
        r = remap(r, b)
    return r, elem, b, energy, np.array(force, dtype=np.float64).reshape(-1, 3)


def to_normalized(r, box, conv_length):
    """Structure::toNormalizedUnits for positions and box."""
    r = r * conv_length
    if box is not None:
        box = box * conv_length
    return r, box
