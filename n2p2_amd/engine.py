"""ctypes binding of the HIP engine (include/hdnnp.h) with a Mode-like API.

Class and method names follow the reference's libnnp call sequence
(Mode.h:297-386, Structure.h:188-190) so that parity tests read like the
reference's own tests:

    m = Model.load("input.nn", "scaling.data", "weights.%03zu.data")
    ctx = Context(m)
    ctx.set_structure(r, elem, box)
    ctx.calculate_neighbor_list(m.max_cutoff)        # Structure::calculateNeighborList
    ctx.calculate_symmetry_function_groups(True)     # Mode::calculateSymmetryFunctionGroups
    ctx.calculate_atomic_neural_networks(True)       # Mode::calculateAtomicNeuralNetworks
    e = ctx.calculate_energy()                       # Mode::calculateEnergy
    ctx.calculate_forces()                           # Mode::calculateForces
    f = ctx.forces()

Errors raise RuntimeError with the engine's "ERROR: ..." message, like the
reference's std::runtime_error.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhdnnp.so")

FORCES_REFERENCE = 0
FORCES_CONSISTENT = 1
PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_MIXED16 = 2
PRECISION_MIXED_BF16 = 3

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)

# (name, restype, argtypes) of every exported function in include/hdnnp.h.
_SIGNATURES = [
    ("hdnnp_last_error", ctypes.c_char_p, []),
    ("hdnnp_model_load", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("hdnnp_model_free", None, [ctypes.c_void_p]),
    ("hdnnp_model_num_elements", ctypes.c_int, [ctypes.c_void_p]),
    ("hdnnp_model_atomic_number", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_model_num_sf", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_model_max_cutoff", ctypes.c_double, [ctypes.c_void_p]),
    ("hdnnp_model_element_index", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("hdnnp_model_normalization", ctypes.c_int,
     [ctypes.c_void_p, _c_double_p, _c_double_p, _c_double_p]),
    ("hdnnp_model_energy_offset", ctypes.c_double, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_model_sf_table", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_double_p]),
    ("hdnnp_model_set_scaling", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _c_double_p]),
    ("hdnnp_ctx_create", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("hdnnp_ctx_destroy", None, [ctypes.c_void_p]),
    ("hdnnp_ctx_set_precision", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_ctx_set_specialized", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_ctx_specialized_active", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    ("hdnnp_set_neighbor_list", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, _c_int32_p, _c_int32_p, _c_int64_p,
      _c_int32_p, _c_int32_p, _c_double_p]),
    ("hdnnp_get_scattered_forces", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, _c_double_p]),
    ("hdnnp_ctx_update_model", ctypes.c_int, [ctypes.c_void_p]),
    ("hdnnp_set_structure", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, _c_double_p, _c_int32_p, _c_double_p]),
    ("hdnnp_set_structures", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, _c_int64_p, _c_double_p, _c_int32_p,
      _c_double_p, _c_int32_p]),
    ("hdnnp_num_structures", ctypes.c_int64, [ctypes.c_void_p]),
    ("hdnnp_get_structure_energies", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("hdnnp_get_structure_pbc", ctypes.c_int, [ctypes.c_void_p, _c_int32_p]),
    ("hdnnp_set_positions", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("hdnnp_set_positions_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("hdnnp_neighbor_list", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("hdnnp_symmetry_functions", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_atomic_networks", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_energy", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("hdnnp_forces", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_evaluate", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_double_p]),
    ("hdnnp_num_atoms", ctypes.c_int64, [ctypes.c_void_p]),
    ("hdnnp_num_neighbor_entries", ctypes.c_int64, [ctypes.c_void_p]),
    ("hdnnp_get_pbc", ctypes.c_int, [ctypes.c_void_p, _c_int32_p]),
    ("hdnnp_get_neighbors", ctypes.c_int,
     [ctypes.c_void_p, _c_int64_p, _c_int32_p, _c_int32_p, _c_double_p, _c_double_p]),
    ("hdnnp_get_G", ctypes.c_int, [ctypes.c_void_p, _c_double_p, ctypes.c_int]),
    ("hdnnp_get_dEdG", ctypes.c_int, [ctypes.c_void_p, _c_double_p, ctypes.c_int]),
    ("hdnnp_get_atomic_energies", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("hdnnp_get_forces", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("hdnnp_forces_device", ctypes.c_void_p, [ctypes.c_void_p]),
    ("hdnnp_get_pair_forces", ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    ("hdnnp_md_init", ctypes.c_int,
     [ctypes.c_void_p, _c_double_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
      _c_double_p, ctypes.c_int, ctypes.c_int]),
    ("hdnnp_md_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _c_double_p, _c_double_p]),
    ("hdnnp_md_get_state", ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    ("hdnnp_md_rebuilds", ctypes.c_int64, [ctypes.c_void_p]),
    ("hdnnp_md_partner_misses", ctypes.c_int64, [ctypes.c_void_p]),
    ("hdnnp_mts_init", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, _c_double_p, ctypes.c_double, ctypes.c_int,
      ctypes.c_double, ctypes.c_double, _c_double_p, ctypes.c_double, ctypes.c_int]),
    ("hdnnp_mts_run", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, _c_double_p, _c_double_p, _c_double_p, _c_int32_p]),
    ("hdnnp_num_weights", ctypes.c_int64, [ctypes.c_void_p]),
    ("hdnnp_get_weights", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("hdnnp_set_weights", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("hdnnp_train_gradient", ctypes.c_int,
     [ctypes.c_void_p, _c_double_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
      _c_double_p, _c_double_p]),
    ("hdnnp_train_adam", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    ("hdnnp_train_reset", ctypes.c_int, [ctypes.c_void_p]),
    ("hdnnp_kalman_init", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    ("hdnnp_kalman_update", ctypes.c_int, [ctypes.c_void_p, _c_double_p, ctypes.c_double]),
    ("hdnnp_energy_weight_jacobian", ctypes.c_int,
     [ctypes.c_void_p, _c_double_p, ctypes.c_int]),
    ("hdnnp_ctx_set_pair_record", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_ctx_pair_record_active", ctypes.c_int, [ctypes.c_void_p, _c_int32_p]),
    ("hdnnp_ctx_autotune", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_double_p,
      ctypes.c_int, _c_int32_p]),
    ("hdnnp_ctx_tuning", ctypes.c_int, [ctypes.c_void_p, _c_int32_p]),
    ("hdnnp_ctx_set_tuning", ctypes.c_int, [ctypes.c_void_p, _c_int32_p]),
    ("hdnnp_ctx_kernel_name", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    ("hdnnp_get_scattered_forces_noalloc", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int64, _c_double_p]),
    ("hdnnp_extrapolation_warnings", ctypes.c_int64, [ctypes.c_void_p, _c_int32_p]),
    ("hdnnp_extrapolation_warning_list", ctypes.c_int64,
     [ctypes.c_void_p, ctypes.c_int64, _c_int32_p, _c_int32_p, _c_double_p,
      _c_double_p, _c_double_p]),
    ("hdnnp_sf_statistics", ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int, _c_double_p]),
    ("hdnnp_sf_statistics_raw", ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int, _c_double_p]),
    ("hdnnp_profile_enable", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hdnnp_profile_avg_ms", ctypes.c_double, [ctypes.c_void_p, ctypes.c_char_p, _c_int64_p]),
    ("hdnnp_profile_reset", ctypes.c_int, [ctypes.c_void_p]),
    ("hdnnp_work_counts", ctypes.c_int, [ctypes.c_void_p, _c_int64_p]),
]

EXPORTED = [s[0] for s in _SIGNATURES]

_lib = None


def load_library(path=None):
    """Loads libhdnnp.so (raises if it was not built).  The environment
    variable N2P2_AMD_LIBRARY names another build of the same library (for
    A/B timing of kernel variants)."""
    global _lib
    if _lib is not None:
        return _lib
    if path is None:
        path = os.environ.get("N2P2_AMD_LIBRARY", LIB_PATH)
    if not os.path.exists(path):
        raise ImportError(
            "n2p2_amd: HIP engine library %s is missing; build it with "
            "`make -C n2p2_amd` (or __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _check(rc):
    if rc != 0:
        msg = _lib.hdnnp_last_error().decode(errors="replace")
        raise RuntimeError("hdnnp error %d: %s" % (rc, msg.strip()))


class Model:
    """Settings + scaling + weights (Mode setup sequence)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def load(cls, settings, scaling, weights_fmt):
        lib = load_library()
        h = ctypes.c_void_p()
        sc = scaling.encode() if scaling else None
        _check(lib.hdnnp_model_load(settings.encode(), sc, weights_fmt.encode(),
                                    ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load_dir(cls, d, scaling=True):
        return cls.load(os.path.join(d, "input.nn"),
                        os.path.join(d, "scaling.data") if scaling else None,
                        os.path.join(d, "weights.%03zu.data"))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hdnnp_model_free(self._h)
            self._h = None

    @property
    def num_elements(self):
        return _lib.hdnnp_model_num_elements(self._h)

    @property
    def max_cutoff(self):
        return _lib.hdnnp_model_max_cutoff(self._h)

    def num_sf(self, e):
        return _lib.hdnnp_model_num_sf(self._h, e)

    def atomic_number(self, e):
        return _lib.hdnnp_model_atomic_number(self._h, e)

    def element_index(self, sym):
        return _lib.hdnnp_model_element_index(self._h, sym.encode())

    def normalization(self):
        me, ce, cl = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        used = _lib.hdnnp_model_normalization(self._h, ctypes.byref(me),
                                              ctypes.byref(ce), ctypes.byref(cl))
        return bool(used), me.value, ce.value, cl.value

    def energy_offset(self, e):
        return _lib.hdnnp_model_energy_offset(self._h, e)

    def sf_table(self, e):
        out = np.zeros((self.num_sf(e), 12))
        _check(_lib.hdnnp_model_sf_table(self._h, e, _ptr(out, ctypes.c_double)))
        return out

    def set_scaling(self, e, scale_type, smin, smax, stats):
        stats = np.ascontiguousarray(stats, dtype=np.float64)
        _check(_lib.hdnnp_model_set_scaling(self._h, e, scale_type, smin, smax,
                                            _ptr(stats, ctypes.c_double)))


class Context:
    """One engine context (device buffers + stream), like one Mode object."""

    def __init__(self, model, device=0):
        lib = load_library()
        self.model = model
        h = ctypes.c_void_p()
        _check(lib.hdnnp_ctx_create(model._h, device, ctypes.byref(h)))
        self._h = h
        self.n = 0
        self.maxnsf = max(model.num_sf(e) for e in range(model.num_elements))

    def close(self):
        if getattr(self, "_h", None):
            _lib.hdnnp_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_precision(self, p):
        _check(_lib.hdnnp_ctx_set_precision(self._h, p))

    def set_specialized(self, on):
        """Model-specialised kernels (default) or the generic ones."""
        _check(_lib.hdnnp_ctx_set_specialized(self._h, 1 if on else 0))

    def specialized_active(self):
        """True when the specialised kernels of the current precision are
        enabled and compiled (compiles them if needed)."""
        v = ctypes.c_int(0)
        _check(_lib.hdnnp_ctx_specialized_active(self._h, ctypes.byref(v)))
        return bool(v.value)

    def set_pair_record(self, on):
        """Mixed paths: the SF kernel records the passing neighbour pairs
        and the force kernel replays them (default), or both enumerate."""
        _check(_lib.hdnnp_ctx_set_pair_record(self._h, 1 if on else 0))

    def pair_record_active(self):
        """True when SF evaluations now record the pairs."""
        v = ctypes.c_int32(0)
        _check(_lib.hdnnp_ctx_pair_record_active(self._h, ctypes.byref(v)))
        return bool(v.value)

    def update_model(self):
        _check(_lib.hdnnp_ctx_update_model(self._h))

    TUNE_COLUMNS = ("chunk", "sf_waves", "force_waves", "nn_tiles", "sf_ms",
                    "force_ms", "nn_ms", "chosen")

    def autotune(self, force_mode=FORCES_CONSISTENT, sweep=False, reps=3):
        """Times the kernels' launch forms on the loaded structure and list
        and applies the fastest (or, unless sweep, a choice cached in this
        process for the same model shape and precision).  Returns the sweep
        rows as a float array with columns TUNE_COLUMNS (empty when a cached
        choice was applied)."""
        cap = 64
        table = np.zeros((cap, len(self.TUNE_COLUMNS)), dtype=np.float64)
        n = ctypes.c_int32(0)
        _check(_lib.hdnnp_ctx_autotune(self._h, int(force_mode),
                                       1 if sweep else 0, int(reps),
                                       _ptr(table, ctypes.c_double), cap,
                                       ctypes.byref(n)))
        return table[:min(n.value, cap)].copy()

    def tuning(self):
        """(member chunk, SF waves, force waves, network tiles)."""
        v = np.zeros(4, dtype=np.int32)
        _check(_lib.hdnnp_ctx_tuning(self._h, _ptr(v, ctypes.c_int32)))
        return tuple(int(x) for x in v)

    def kernel_name(self, which):
        """Name of the SF ("sf") or force ("force") kernel as profilers
        list it, at the context's precision and tuning."""
        buf = ctypes.create_string_buffer(128)
        _check(_lib.hdnnp_ctx_kernel_name(self._h,
                                          {"sf": 0, "force": 1}[which], buf,
                                          128))
        return buf.value.decode()

    def set_tuning(self, chunk=16, sf_waves=-1, force_waves=-1, nn_tiles=1):
        v = np.array([chunk, sf_waves, force_waves, nn_tiles], dtype=np.int32)
        _check(_lib.hdnnp_ctx_set_tuning(self._h, _ptr(v, ctypes.c_int32)))

    def set_structure(self, r, elem, box=None):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
        elem = np.ascontiguousarray(elem, dtype=np.int32)
        self._r, self._e = r, elem
        bp = None
        if box is not None:
            self._box = np.ascontiguousarray(box, dtype=np.float64).reshape(3, 3)
            bp = _ptr(self._box, ctypes.c_double)
        _check(_lib.hdnnp_set_structure(self._h, r.shape[0], _ptr(r, ctypes.c_double),
                                        _ptr(elem, ctypes.c_int32), bp))
        self.n = r.shape[0]

    def set_structures(self, structures):
        """A batch of structures [(r, elem, box or None), ...] evaluated
        together (the nnp-dataset loop over Mode::evaluateNNP).  Results
        are over the concatenated atoms; structure_energies() splits the
        energy."""
        rs, es, boxes, per, off = [], [], [], [], [0]
        for r, e, box in structures:
            r = np.asarray(r, dtype=np.float64).reshape(-1, 3)
            rs.append(r)
            es.append(np.asarray(e, dtype=np.int32))
            boxes.append(np.zeros((3, 3)) if box is None
                         else np.asarray(box, dtype=np.float64).reshape(3, 3))
            per.append(0 if box is None else 1)
            off.append(off[-1] + r.shape[0])
        r = np.ascontiguousarray(np.concatenate(rs))
        elem = np.ascontiguousarray(np.concatenate(es))
        bx = np.ascontiguousarray(np.stack(boxes))
        per = np.ascontiguousarray(per, dtype=np.int32)
        off = np.ascontiguousarray(off, dtype=np.int64)
        self._r, self._e, self._box = r, elem, bx
        _check(_lib.hdnnp_set_structures(
            self._h, len(rs), _ptr(off, ctypes.c_int64),
            _ptr(r, ctypes.c_double), _ptr(elem, ctypes.c_int32),
            _ptr(bx, ctypes.c_double), _ptr(per, ctypes.c_int32)))
        self.n = r.shape[0]
        self.struct_offsets = off

    def num_structures(self):
        return _lib.hdnnp_num_structures(self._h)

    def structure_energies(self):
        e = np.zeros(self.num_structures())
        _check(_lib.hdnnp_get_structure_energies(self._h, _ptr(e, ctypes.c_double)))
        return e

    def structure_pbc(self):
        p = np.zeros((self.num_structures(), 3), dtype=np.int32)
        _check(_lib.hdnnp_get_structure_pbc(self._h, _ptr(p, ctypes.c_int32)))
        return p

    # --- MD (velocity Verlet, Verlet-skin list, HIP graphs)
    FTOA_EV_A_AMU_FS = 9.648533212e-3

    def md_init(self, masses, dt, skin=1.0, v=None, ftoa=FTOA_EV_A_AMU_FS,
                force_mode=FORCES_CONSISTENT, graph=True):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        vp = None
        if v is not None:
            self._v0 = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
            vp = _ptr(self._v0, ctypes.c_double)
        _check(_lib.hdnnp_md_init(self._h, _ptr(m, ctypes.c_double), dt, skin,
                                  ftoa, vp, force_mode, int(graph)))

    def md_run(self, nsteps):
        """nsteps steps; returns (epot, ekin) per step."""
        ep = np.zeros(nsteps)
        ek = np.zeros(nsteps)
        _check(_lib.hdnnp_md_run(self._h, nsteps, _ptr(ep, ctypes.c_double),
                                 _ptr(ek, ctypes.c_double)))
        return ep, ek

    def md_state(self):
        r = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        _check(_lib.hdnnp_md_get_state(self._h, _ptr(r, ctypes.c_double),
                                       _ptr(v, ctypes.c_double)))
        return r, v

    def md_rebuilds(self):
        return _lib.hdnnp_md_rebuilds(self._h)

    def md_partner_misses(self):
        return _lib.hdnnp_md_partner_misses(self._h)

    # --- multiple-time-step MD (this context = the fast surrogate)
    def mts_init(self, full, masses, dt_outer=1.0, n_inner=4, skin=1.0,
                 v=None, rmse_gate=5e-3, ftoa=FTOA_EV_A_AMU_FS, graph=True):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        vp = None
        if v is not None:
            self._v0 = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
            vp = _ptr(self._v0, ctypes.c_double)
        _check(_lib.hdnnp_mts_init(full._h, self._h, _ptr(m, ctypes.c_double),
                                   dt_outer, n_inner, skin, ftoa, vp,
                                   rmse_gate, int(graph)))

    def mts_run(self, nouter):
        """nouter outer steps; returns (epot, ekin, rmse, mode) per step."""
        ep, ek, rm = np.zeros(nouter), np.zeros(nouter), np.zeros(nouter)
        md = np.zeros(nouter, dtype=np.int32)
        _check(_lib.hdnnp_mts_run(self._h, nouter, _ptr(ep, ctypes.c_double),
                                  _ptr(ek, ctypes.c_double),
                                  _ptr(rm, ctypes.c_double),
                                  _ptr(md, ctypes.c_int32)))
        return ep, ek, rm, md

    # --- training of a radial-only model
    def num_weights(self):
        return _lib.hdnnp_num_weights(self._h)

    def get_weights(self):
        w = np.zeros(self.num_weights())
        _check(_lib.hdnnp_get_weights(self._h, _ptr(w, ctypes.c_double)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.num_weights()
        _check(_lib.hdnnp_set_weights(self._h, _ptr(w, ctypes.c_double)))

    def train_gradient(self, f_ref, e_ref, w_e=1.0, w_f=1.0, want_grad=True):
        """Gradient of w_e ((E - e_ref)/N)^2 + w_f/(3N) sum |F - f_ref|^2 at
        the loaded structure; returns (grad or None, (loss_e, loss_f))."""
        f = np.ascontiguousarray(f_ref, dtype=np.float64).reshape(-1, 3)
        g = np.zeros(self.num_weights()) if want_grad else None
        loss = np.zeros(2)
        _check(_lib.hdnnp_train_gradient(
            self._h, _ptr(f, ctypes.c_double), e_ref, w_e, w_f,
            _ptr(g, ctypes.c_double) if g is not None else None,
            _ptr(loss, ctypes.c_double)))
        return g, (float(loss[0]), float(loss[1]))

    def train_adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(_lib.hdnnp_train_adam(self._h, lr, beta1, beta2, eps))

    def train_reset(self):
        _check(_lib.hdnnp_train_reset(self._h))

    def kalman_init(self, epsilon=1e-2, q0=0.01, qtau=2.302, qmin=1e-6,
                    eta0=0.01, etatau=2.302, etamax=1.0):
        _check(_lib.hdnnp_kalman_init(self._h, epsilon, q0, qtau, qmin, eta0,
                                      etatau, etamax))

    def kalman_update(self, h, xi):
        """One scalar-observation Kalman update of the weights: Jacobian h
        (flat weight order), innovation xi."""
        h = np.ascontiguousarray(h, dtype=np.float64)
        _check(_lib.hdnnp_kalman_update(self._h, _ptr(h, ctypes.c_double),
                                        float(xi)))

    def energy_weight_jacobian(self, file_order=True):
        """dE_s/dw for every loaded structure s: (nstruct, nweights)."""
        ns = max(self.num_structures(), 1)
        J = np.zeros((ns, self.num_weights()))
        _check(_lib.hdnnp_energy_weight_jacobian(
            self._h, _ptr(J, ctypes.c_double), int(file_order)))
        return J

    def set_neighbor_list(self, centre_elem, offsets, nbr_index, nbr_elem,
                          dr, centre_index=None):
        """Caller-given neighbour lists (InterfaceLammps::addNeighbor):
        centres with element centre_elem, entries offsets[i]:offsets[i+1]
        with the caller's neighbour index, its element and r_i - r_j."""
        ce = np.ascontiguousarray(centre_elem, dtype=np.int32)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        ni = np.ascontiguousarray(nbr_index, dtype=np.int32)
        ne = np.ascontiguousarray(nbr_elem, dtype=np.int32)
        d = np.ascontiguousarray(dr, dtype=np.float64).reshape(-1, 3)
        ci = (None if centre_index is None
              else np.ascontiguousarray(centre_index, dtype=np.int32))
        i32, i64, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        _check(_lib.hdnnp_set_neighbor_list(
            self._h, len(ce), _ptr(ci, i32) if ci is not None else None,
            _ptr(ce, i32), _ptr(off, i64), _ptr(ni, i32), _ptr(ne, i32),
            _ptr(d, f64)))
        self.n = len(ce)

    def scattered_forces(self, nout):
        """Forces on the caller's nout atoms from a caller-given list."""
        f = np.zeros((nout, 3))
        _check(_lib.hdnnp_get_scattered_forces(self._h, nout,
                                               _ptr(f, ctypes.c_double)))
        return f

    def scattered_forces_noalloc(self, nout):
        """scattered_forces through the entry point that allocates nothing
        (staging in the library's fixed device array)."""
        f = np.zeros((nout, 3))
        _check(_lib.hdnnp_get_scattered_forces_noalloc(
            self._h, nout, _ptr(f, ctypes.c_double)))
        return f

    def set_positions(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
        _check(_lib.hdnnp_set_positions(self._h, _ptr(r, ctypes.c_double)))

    def set_positions_device(self, ptr):
        _check(_lib.hdnnp_set_positions_device(self._h, ctypes.c_void_p(ptr)))

    # --- Mode / Structure compute sequence
    def calculate_neighbor_list(self, cutoff):
        _check(_lib.hdnnp_neighbor_list(self._h, cutoff))

    def calculate_symmetry_function_groups(self, derivatives=True):
        _check(_lib.hdnnp_symmetry_functions(self._h, int(derivatives)))

    def calculate_atomic_neural_networks(self, derivatives=True):
        _check(_lib.hdnnp_atomic_networks(self._h, int(derivatives)))

    def calculate_energy(self):
        e = ctypes.c_double()
        _check(_lib.hdnnp_energy(self._h, ctypes.byref(e)))
        return e.value

    def calculate_forces(self, mode=FORCES_REFERENCE):
        _check(_lib.hdnnp_forces(self._h, mode))

    def evaluate(self, force_mode=FORCES_REFERENCE):
        """Mode::evaluateNNP; force_mode < 0 skips forces."""
        e = ctypes.c_double()
        _check(_lib.hdnnp_evaluate(self._h, force_mode, ctypes.byref(e)))
        return e.value

    # --- results
    def num_atoms(self):
        """Atoms (or centres) of the loaded structure; 0 when none is."""
        return int(_lib.hdnnp_num_atoms(self._h))

    def num_neighbor_entries(self):
        return _lib.hdnnp_num_neighbor_entries(self._h)

    def pbc(self):
        p = np.zeros(3, dtype=np.int32)
        _check(_lib.hdnnp_get_pbc(self._h, _ptr(p, ctypes.c_int32)))
        return p

    def neighbors(self):
        m = self.num_neighbor_entries()
        off = np.zeros(self.n + 1, dtype=np.int64)
        idx = np.zeros(m, dtype=np.int32)
        img = np.zeros((m, 3), dtype=np.int32)
        d = np.zeros(m)
        dr = np.zeros((m, 3))
        _check(_lib.hdnnp_get_neighbors(self._h, _ptr(off, ctypes.c_int64),
                                        _ptr(idx, ctypes.c_int32), _ptr(img, ctypes.c_int32),
                                        _ptr(d, ctypes.c_double), _ptr(dr, ctypes.c_double)))
        return off, idx, img, d, dr

    def G(self):
        out = np.zeros((self.n, self.maxnsf))
        _check(_lib.hdnnp_get_G(self._h, _ptr(out, ctypes.c_double), self.maxnsf))
        return out

    def dEdG(self):
        out = np.zeros((self.n, self.maxnsf))
        _check(_lib.hdnnp_get_dEdG(self._h, _ptr(out, ctypes.c_double), self.maxnsf))
        return out

    def atomic_energies(self):
        out = np.zeros(self.n)
        _check(_lib.hdnnp_get_atomic_energies(self._h, _ptr(out, ctypes.c_double)))
        return out

    def forces(self):
        out = np.zeros((self.n, 3))
        _check(_lib.hdnnp_get_forces(self._h, _ptr(out, ctypes.c_double)))
        return out

    def pair_forces(self):
        """(self forces n x 3, pair forces M x 3) as the reference's
        Atom::calculateSelfForceShort / calculatePairForceShort."""
        sf = np.zeros((self.n, 3))
        pf = np.zeros((self.num_neighbor_entries(), 3))
        _check(_lib.hdnnp_get_pair_forces(self._h, _ptr(sf, ctypes.c_double),
                                          _ptr(pf, ctypes.c_double)))
        return sf, pf

    def forces_device_ptr(self):
        return _lib.hdnnp_forces_device(self._h)

    def extrapolation_warnings(self):
        per = np.zeros(self.n, dtype=np.int32)
        tot = _lib.hdnnp_extrapolation_warnings(self._h, _ptr(per, ctypes.c_int32))
        if tot < 0:
            _check(int(tot))
        return tot, per

    def extrapolation_warning_list(self):
        """Records (atom, sf, value, gmin, gmax) of every extrapolation
        warning of the last SF evaluation, in the reference's order."""
        tot = _lib.hdnnp_extrapolation_warning_list(self._h, 0, None, None,
                                                    None, None, None)
        if tot < 0:
            _check(int(tot))
        a = np.zeros(tot, dtype=np.int32)
        k = np.zeros(tot, dtype=np.int32)
        v, lo, hi = np.zeros(tot), np.zeros(tot), np.zeros(tot)
        i32, f64 = ctypes.c_int32, ctypes.c_double
        _lib.hdnnp_extrapolation_warning_list(self._h, tot, _ptr(a, i32),
                                              _ptr(k, i32), _ptr(v, f64),
                                              _ptr(lo, f64), _ptr(hi, f64))
        return a, k, v, lo, hi

    def sf_statistics(self, e):
        out = np.zeros((self.model.num_sf(e), 4))
        na = _lib.hdnnp_sf_statistics(self._h, e, _ptr(out, ctypes.c_double))
        if na < 0:
            _check(int(na))
        return na, out

    def sf_statistics_raw(self, e):
        """(count, [nsf x 4]: min, max, sum, sum of squares) of element e."""
        out = np.zeros((self.model.num_sf(e), 4))
        na = _lib.hdnnp_sf_statistics_raw(self._h, e, _ptr(out, ctypes.c_double))
        if na < 0:
            _check(int(na))
        return na, out

    def work_counts(self):
        """hdnnp_work_counts: 10 exact counts of pairs, triplets and member
        evaluations for the loaded structure (see include/hdnnp.h)."""
        c = np.zeros(10, dtype=np.int64)
        _check(_lib.hdnnp_work_counts(self._h, _ptr(c, ctypes.c_int64)))
        return c

    # --- profiling
    def profile(self, enable=True):
        _check(_lib.hdnnp_profile_enable(self._h, int(enable)))

    def profile_reset(self):
        _check(_lib.hdnnp_profile_reset(self._h))

    def profile_avg_ms(self, name):
        calls = ctypes.c_int64()
        ms = _lib.hdnnp_profile_avg_ms(self._h, name.encode(), ctypes.byref(calls))
        return ms, calls.value
