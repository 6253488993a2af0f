"""A LAMMPS-like domain for the caller-given neighbour-list path: the atoms
of a periodic box as local atoms, their periodic images within the cutoff as
explicit ghost atoms with open boundaries, and full neighbour lists of the
local atoms over local and ghost atoms (what pair_hdnnp.cpp:81-114 hands to
InterfaceLammps::addNeighbor, with dr = x_i - x_j).  Folding the ghost forces
back onto their owners gives the forces of the periodic system including the
periodic self-images (HDNNP_FORCES_CONSISTENT)."""
import numpy as np


def ghost_domain(r, elem, box, rc):
    """Returns (x, el, owner, nlocal): local atoms first, then every
    periodic image within rc of the box's atoms."""
    r = np.asarray(r, dtype=float)
    box = np.asarray(box, dtype=float)
    n = len(r)
    # Enough images in every direction for a cutoff rc.
    inv = np.linalg.inv(box)
    widths = 1.0 / np.linalg.norm(inv, axis=0)
    p = np.ceil(rc / widths).astype(int)
    xs, es, own = [r], [np.asarray(elem)], [np.arange(n)]
    frac = r @ inv
    for a in range(-p[0], p[0] + 1):
        for b in range(-p[1], p[1] + 1):
            for c in range(-p[2], p[2] + 1):
                if a == b == c == 0:
                    continue
                shift = a * box[0] + b * box[1] + c * box[2]
                img = r + shift
                # keep images within rc of the box (fractional margin)
                f = frac + np.array([a, b, c])
                m = np.ones(n, dtype=bool)
                for k in range(3):
                    lo = -rc / widths[k]
                    hi = 1.0 + rc / widths[k]
                    m &= (f[:, k] >= lo) & (f[:, k] <= hi)
                xs.append(img[m])
                es.append(np.asarray(elem)[m])
                own.append(np.arange(n)[m])
    return (np.concatenate(xs), np.concatenate(es), np.concatenate(own), n)


def full_lists(x, nlocal, rc):
    """Full neighbour lists of the local atoms (LAMMPS order: ascending
    index over local and ghost atoms): offsets, neighbour indices and
    dr = x_i - x_j."""
    off = [0]
    idx, dr = [], []
    rc2 = rc * rc
    for i in range(nlocal):
        d = x[i] - x
        d2 = np.einsum("ij,ij->i", d, d)
        js = np.nonzero((d2 <= rc2) & (d2 > 0.0))[0]
        idx.append(js)
        dr.append(d[js])
        off.append(off[-1] + len(js))
    return (np.array(off, dtype=np.int64), np.concatenate(idx).astype(np.int32),
            np.concatenate(dr))
