"""ctypes binding of the oracle (oracle/liboracle.so) -- test infrastructure.

The oracle is the CPU restatement of the reference path used as the checker
for the HIP engine.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_V = ctypes.c_void_p

_SIG = [
    ("orc_last_error", ctypes.c_char_p, []),
    ("orc_model_load", _V, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    ("orc_model_free", None, [_V]),
    ("orc_model_num_elements", ctypes.c_int, [_V]),
    ("orc_model_atomic_number", ctypes.c_int, [_V, ctypes.c_int]),
    ("orc_model_num_sf", ctypes.c_int, [_V, ctypes.c_int]),
    ("orc_model_max_cutoff", ctypes.c_double, [_V]),
    ("orc_model_normalize", ctypes.c_int, [_V, _D, _D, _D]),
    ("orc_model_sf_table", ctypes.c_int, [_V, ctypes.c_int, _D]),
    ("orc_model_groups", ctypes.c_int, [_V, ctypes.c_int, _I32, ctypes.c_int]),
    ("orc_struct_create", _V, [_V, ctypes.c_int64, _D, _I32, _D]),
    ("orc_struct_read", _V, [_V, ctypes.c_char_p]),
    ("orc_struct_free", None, [_V]),
    ("orc_struct_num_atoms", ctypes.c_int64, [_V]),
    ("orc_struct_remap", None, [_V]),
    ("orc_struct_get_positions", None, [_V, _D]),
    ("orc_struct_get_elements", None, [_V, _I32]),
    ("orc_struct_get_box", None, [_V, _D]),
    ("orc_struct_to_normalized", ctypes.c_int, [_V, _V]),
    ("orc_neighbor_list", ctypes.c_int64, [_V, ctypes.c_double, ctypes.c_int]),
    ("orc_get_pbc", None, [_V, _I32]),
    ("orc_get_neighbors", None, [_V, _I64, _I32, _D, _D]),
    ("orc_symmetry_functions", ctypes.c_int, [_V, _V, ctypes.c_int]),
    ("orc_networks", ctypes.c_int, [_V, _V, ctypes.c_int]),
    ("orc_energy", ctypes.c_double, [_V]),
    ("orc_forces", ctypes.c_int, [_V, _V, ctypes.c_int]),
    ("orc_get_G", ctypes.c_int, [_V, _D, ctypes.c_int]),
    ("orc_get_dEdG", ctypes.c_int, [_V, _D, ctypes.c_int]),
    ("orc_get_atomic_energies", None, [_V, _D]),
    ("orc_get_forces", None, [_V, _D]),
    ("orc_get_dGdr_self", ctypes.c_int, [_V, ctypes.c_int64, _D]),
    ("orc_get_dGdr_neighbor", ctypes.c_int, [_V, ctypes.c_int64, ctypes.c_int64, _D]),
    ("orc_get_pair_forces", ctypes.c_int, [_V, _V, _D, _D]),
    ("orc_predict_file", ctypes.c_int, [_V, ctypes.c_char_p, _D, _D, ctypes.c_int64, _I64]),
    ("orc_evaluate", ctypes.c_double, [_V, _V, ctypes.c_int, ctypes.c_int]),
    ("orc_energy_only", ctypes.c_double, [_V, _V, ctypes.c_int]),
    ("orc_sf_statistics", ctypes.c_int, [_V, _V, ctypes.c_int, _D]),
    ("orc_extrapolation_warnings", ctypes.c_int64,
     [_V, _V, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
      _D, ctypes.c_int64]),
    ("orc_numeric_forces", ctypes.c_int,
     [_V, _V, ctypes.c_double, _I64, ctypes.c_int64, ctypes.c_int, _D]),
    ("orc_num_threads", ctypes.c_int, []),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-C", ORACLE_DIR, "-s"])
        L = ctypes.CDLL(LIB)
        for n, r, a in _SIG:
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _err():
    return lib().orc_last_error().decode()


class OModel:
    def __init__(self, settings, scaling, weights_fmt):
        L = lib()
        h = L.orc_model_load(settings.encode(),
                             scaling.encode() if scaling else None,
                             weights_fmt.encode())
        if not h:
            raise RuntimeError(_err())
        self.h = ctypes.c_void_p(h)

    @classmethod
    def from_dir(cls, d, scaling=True):
        return cls(os.path.join(d, "input.nn"),
                   os.path.join(d, "scaling.data") if scaling else None,
                   os.path.join(d, "weights.%03zu.data"))

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.orc_model_free(self.h)
            self.h = None

    @property
    def num_elements(self):
        return lib().orc_model_num_elements(self.h)

    @property
    def max_cutoff(self):
        return lib().orc_model_max_cutoff(self.h)

    def num_sf(self, e):
        return lib().orc_model_num_sf(self.h, e)

    def sf_table(self, e):
        out = np.zeros((self.num_sf(e), 12))
        lib().orc_model_sf_table(self.h, e, _p(out))
        return out

    def groups(self, e):
        n = lib().orc_model_groups(self.h, e, None, 0)
        buf = np.zeros(4096, dtype=np.int32)
        lib().orc_model_groups(self.h, e, _p(buf, ctypes.c_int32), 4096)
        out, pos = [], 0
        for _ in range(n):
            t, e1, e2, nm = buf[pos:pos + 4]
            out.append((int(t), int(e1), int(e2), [int(x) for x in buf[pos + 4:pos + 4 + nm]]))
            pos += 4 + nm
        return out

    def normalization(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        used = lib().orc_model_normalize(self.h, ctypes.byref(a), ctypes.byref(b),
                                         ctypes.byref(c))
        return bool(used), a.value, b.value, c.value

    def predict_file(self, path):
        cap = 1 << 20
        f = np.zeros((cap, 3))
        e = ctypes.c_double()
        n = ctypes.c_int64()
        rc = lib().orc_predict_file(self.h, path.encode(), ctypes.byref(e), _p(f), cap,
                                    ctypes.byref(n))
        if rc:
            raise RuntimeError(_err())
        return e.value, f[:n.value].copy()


class OStruct:
    def __init__(self, model, r=None, elem=None, box=None, path=None):
        L = lib()
        self.model = model
        if path is not None:
            h = L.orc_struct_read(model.h, path.encode())
        else:
            r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
            elem = np.ascontiguousarray(elem, dtype=np.int32)
            bp = None
            if box is not None:
                self._box = np.ascontiguousarray(box, dtype=np.float64).reshape(3, 3)
                bp = _p(self._box)
            h = L.orc_struct_create(model.h, r.shape[0], _p(r), _p(elem, ctypes.c_int32), bp)
        if not h:
            raise RuntimeError(_err())
        self.h = ctypes.c_void_p(h)
        self.n = L.orc_struct_num_atoms(self.h)
        self.maxnsf = max(model.num_sf(e) for e in range(model.num_elements))

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            _lib.orc_struct_free(self.h)
            self.h = None

    def positions(self):
        r = np.zeros((self.n, 3))
        lib().orc_struct_get_positions(self.h, _p(r))
        return r

    def elements(self):
        e = np.zeros(self.n, dtype=np.int32)
        lib().orc_struct_get_elements(self.h, _p(e, ctypes.c_int32))
        return e

    def box(self):
        b = np.zeros((3, 3))
        lib().orc_struct_get_box(self.h, _p(b))
        return b

    def remap(self):
        lib().orc_struct_remap(self.h)

    def to_normalized(self):
        return lib().orc_struct_to_normalized(self.h, self.model.h)

    def neighbor_list(self, rc, cells=False):
        return lib().orc_neighbor_list(self.h, rc, int(cells))

    def pbc(self):
        p = np.zeros(3, dtype=np.int32)
        lib().orc_get_pbc(self.h, _p(p, ctypes.c_int32))
        return p

    def neighbors(self):
        off = np.zeros(self.n + 1, dtype=np.int64)
        lib().orc_get_neighbors(self.h, _p(off, ctypes.c_int64), None, None, None)
        m = int(off[-1])
        idx = np.zeros(m, dtype=np.int32)
        d = np.zeros(m)
        dr = np.zeros((m, 3))
        lib().orc_get_neighbors(self.h, _p(off, ctypes.c_int64), _p(idx, ctypes.c_int32),
                                _p(d), _p(dr))
        return off, idx, d, dr

    def symmetry_functions(self, derivatives=True):
        if lib().orc_symmetry_functions(self.h, self.model.h, int(derivatives)):
            raise RuntimeError(_err())

    def networks(self, derivatives=True):
        if lib().orc_networks(self.h, self.model.h, int(derivatives)):
            raise RuntimeError(_err())

    def energy(self):
        return lib().orc_energy(self.h)

    def forces_calc(self, mode=0):
        if lib().orc_forces(self.h, self.model.h, mode):
            raise RuntimeError(_err())

    def G(self):
        out = np.zeros((self.n, self.maxnsf))
        lib().orc_get_G(self.h, _p(out), self.maxnsf)
        return out

    def dEdG(self):
        out = np.zeros((self.n, self.maxnsf))
        lib().orc_get_dEdG(self.h, _p(out), self.maxnsf)
        return out

    def atomic_energies(self):
        out = np.zeros(self.n)
        lib().orc_get_atomic_energies(self.h, _p(out))
        return out

    def forces(self):
        out = np.zeros((self.n, 3))
        lib().orc_get_forces(self.h, _p(out))
        return out

    def pair_forces(self):
        off, _, _, _ = self.neighbors()
        sf = np.zeros((self.n, 3))
        pf = np.zeros((int(off[-1]), 3))
        lib().orc_get_pair_forces(self.h, self.model.h, _p(sf), _p(pf))
        return sf, pf

    def dGdr_self(self, i):
        out = np.zeros((self.maxnsf, 3))
        n = lib().orc_get_dGdr_self(self.h, i, _p(out))
        return out[:n]

    def dGdr_neighbor(self, i, k):
        out = np.zeros((self.maxnsf, 3))
        n = lib().orc_get_dGdr_neighbor(self.h, i, k, _p(out))
        return out[:n]

    def evaluate(self, cells=False, force_mode=0):
        return lib().orc_evaluate(self.h, self.model.h, int(cells), force_mode)

    def energy_only(self, cells=False):
        return lib().orc_energy_only(self.h, self.model.h, int(cells))

    def sf_statistics(self, e):
        out = np.zeros((self.model.num_sf(e), 4))
        n = lib().orc_sf_statistics(self.h, self.model.h, e, _p(out))
        if n < 0:
            raise RuntimeError(_err())
        return n, out

    def extrapolation_warnings(self):
        """(atom, sf, value) records of Element::
        updateSymmetryFunctionStatistics after the SFs were computed."""
        i32 = ctypes.c_int32
        z = np.zeros(1, dtype=np.int32)
        zd = np.zeros(1)
        n = lib().orc_extrapolation_warnings(self.h, self.model.h, _p(z, i32),
                                             _p(z, i32), _p(zd), 0)
        if n < 0:
            raise RuntimeError(_err())
        a = np.zeros(max(n, 1), dtype=np.int32)
        k = np.zeros(max(n, 1), dtype=np.int32)
        v = np.zeros(max(n, 1))
        lib().orc_extrapolation_warnings(self.h, self.model.h, _p(a, i32),
                                         _p(k, i32), _p(v), n)
        return a[:n], k[:n], v[:n]

    def numeric_forces(self, atoms, delta=1e-4, cells=False):
        atoms = np.ascontiguousarray(atoms, dtype=np.int64)
        out = np.zeros((len(atoms), 3))
        if lib().orc_numeric_forces(self.h, self.model.h, delta, _p(atoms, ctypes.c_int64),
                                    len(atoms), int(cells), _p(out)):
            raise RuntimeError(_err())
        return out
