/*
 * hdnnp.h -- C ABI of the MI355X HDNNP energy-and-force engine.
 *
 * This is the drop-in boundary under n2p2's libnnp Mode API.  Every entry
 * point replaces one reference interface, cited as file:line relative to the
 * n2p2 tree (src/libnnp unless stated).  The header-compatible C++ layer in
 * n2p2_amd/nnp (nnp::Mode, nnp::Prediction, nnp::Structure, nnp::Atom, ...,
 * built as libnnp_amd.so) calls these functions; ctypes / cgo / JNI
 * bindings can call them directly.
 *
 * Conventions
 *   - plain C types only; arrays are caller-owned host memory unless the
 *     name ends in _device (then they are device pointers on the context's
 *     GPU);
 *   - every function returning int returns HDNNP_OK (0) on success or a
 *     negative HDNNP_E_* code; hdnnp_last_error() then holds a message that
 *     starts with "ERROR: " like the reference's exceptions;
 *   - a context is single-threaded (like a reference Mode object): calls on
 *     one context must not overlap; distinct contexts are independent;
 *   - all compute calls are synchronous at return unless stated.
 *   - units are the model's internal units (normalised when the settings
 *     file has mean_energy / conv_energy / conv_length, as in Mode).
 */
#ifndef HDNNP_H
#define HDNNP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDNNP_OK            0
#define HDNNP_E_INVALID    -1   /* bad argument / inconsistent call order */
#define HDNNP_E_FILE       -2   /* file missing or malformed */
#define HDNNP_E_SETTINGS   -3   /* settings keyword error */
#define HDNNP_E_DEVICE     -4   /* HIP runtime error */
#define HDNNP_E_CAPACITY   -5   /* a size exceeds a documented limit */
#define HDNNP_E_UNSUPPORTED -6  /* feature outside this engine's scope */

/* Force-gather modes for hdnnp_forces / hdnnp_evaluate. */
#define HDNNP_FORCES_REFERENCE 0  /* Mode::calculateForces: an atom's own
                                     periodic self-image slots are skipped
                                     (Mode.cpp:1885-1887) */
#define HDNNP_FORCES_CONSISTENT 1 /* adds those slots: forces = -dE/dr also
                                     when the cell is smaller than the cutoff */

/* Arithmetic of the descriptor and network kernels. */
#define HDNNP_PRECISION_FP64  0   /* plain path: FP64 throughout */
#define HDNNP_PRECISION_MIXED 1   /* FP32 symmetry-function and force-
                                     contraction arithmetic (per-lane FP32
                                     partial sums combined in FP64), FP32
                                     MFMA network layers; FP64 SF scaling,
                                     biases and activations, energy total
                                     and force accumulation */
#define HDNNP_PRECISION_MIXED16 2 /* as MIXED with FP16 MFMA network layers
                                     (FP32 accumulation) */
#define HDNNP_PRECISION_MIXED_BF16 3 /* as MIXED with BF16 MFMA network
                                        layers (FP32 accumulation) */

typedef struct hdnnp_model hdnnp_model;
typedef struct hdnnp_ctx   hdnnp_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char* hdnnp_last_error(void);

/* ------------------------------------------------------------ model -- */

/* Reads the settings file, the scaling file and one weight file per element
 * and sets the model up exactly as the reference sequence
 *   Mode::initialize; loadSettingsFile; setupGeneric;
 *   setupSymmetryFunctionScaling (or ...ScalingNone if scaling_file is NULL
 *   or ""); setupNeuralNetworkWeights
 * (Mode.cpp:55-236, :692-840, :1158-1478, :2270-2289; Prediction.cpp:34-46).
 * weights_fmt is a printf format with one size_t conversion for the atomic
 * number, e.g. "weights.%03zu.data".  Every symmetry-function type of the
 * reference (2, 3, 9; weighted 12, 13; compact 20-25) and all cutoff,
 * scaling and activation types are supported; unknown SF types, more than
 * 96 compact SFs in one group and type-13 SFs with rs < 0 return
 * HDNNP_E_UNSUPPORTED. */
int  hdnnp_model_load(const char* settings_file, const char* scaling_file,
                      const char* weights_fmt, hdnnp_model** out);
void hdnnp_model_free(hdnnp_model* m);

/* Element count / atomic number of element e (elements sorted by Z,
 * ElementMap.cpp:36-54) / SF count of element e (Element::
 * numSymmetryFunctions) / Mode::getMaxCutoffRadius (Mode.h:468). */
int    hdnnp_model_num_elements(const hdnnp_model* m);
int    hdnnp_model_atomic_number(const hdnnp_model* m, int e);
int    hdnnp_model_num_sf(const hdnnp_model* m, int e);
double hdnnp_model_max_cutoff(const hdnnp_model* m);
/* Element index of a chemical symbol (ElementMap::operator[]), or -1. */
int    hdnnp_model_element_index(const hdnnp_model* m, const char* symbol);
/* Normalisation constants (Mode.cpp:238-299); returns 1 if used, else 0. */
int    hdnnp_model_normalization(const hdnnp_model* m, double* mean_energy,
                                 double* conv_energy, double* conv_length);
/* Element::getAtomicEnergyOffset (atom_energy keyword, Mode.cpp:340-350). */
double hdnnp_model_energy_offset(const hdnnp_model* m, int e);
/* Sorted SF table of element e, 12 doubles per SF in NN-input order:
 * type, e1, e2 (-1 for radial), eta, rs, lambda, zeta, rc, cutoff_type,
 * cutoff_alpha, scaling_factor, settings line number
 * (Element::sortSymmetryFunctions, Element.cpp:154-166). */
int    hdnnp_model_sf_table(const hdnnp_model* m, int e, double* out);
/* Replaces element e's scaling statistics (4 doubles per SF: Gmin, Gmax,
 * Gmean, Gsigma) and recomputes the scaling factors, as
 * SymFnc::setScalingType + SymGrp*::setScalingFactors do when a scaling.data
 * line is read (SymFnc.cpp:52-98, SymGrpBaseExpAng.cpp:45-57).  scale_type
 * is 0 none, 1 scale, 2 center, 3 scale+center, 4 scale-sigma. */
int    hdnnp_model_set_scaling(hdnnp_model* m, int e, int scale_type,
                               double smin, double smax,
                               const double* stats);

/* ---------------------------------------------------------- context -- */

/* One engine context per Mode object: device copies of the model tables,
 * structure and work buffers, one HIP stream.  device = HIP ordinal. */
int  hdnnp_ctx_create(const hdnnp_model* m, int device, hdnnp_ctx** out);
void hdnnp_ctx_destroy(hdnnp_ctx* c);
/* HDNNP_PRECISION_FP64 (default), _MIXED, _MIXED16 or _MIXED_BF16. */
int  hdnnp_ctx_set_precision(hdnnp_ctx* c, int precision);
/* 1 (default unless the environment sets N2P2_AMD_RTC=0): the symmetry-
 * function and force kernels are compiled for this model at first use
 * (member constants folded in, hipRTC) and kept for the process; 0: the
 * generic kernels that read the member tables.  Same results to rounding.
 * No counterpart in the reference (an engine control like the precision). */
int  hdnnp_ctx_set_specialized(hdnnp_ctx* c, int enable);
/* *active = 1 when the specialised kernels for the current precision are
 * enabled and compiled (compiling them now if they were not yet), 0 when
 * the generic kernels run.  No counterpart in the reference. */
int  hdnnp_ctx_specialized_active(hdnnp_ctx* c, int* active);
/* 1 (default unless the environment sets N2P2_AMD_PAIR_RECORD=0): on the
 * mixed-precision paths the symmetry-function kernel records, per atom,
 * the neighbour pairs of every angular group that pass the cutoff tests,
 * and the force kernel of the same evaluation replays them instead of
 * enumerating and testing all pairs again; 0: both kernels enumerate.
 * Same results bit for bit.  No counterpart in the reference (an engine
 * control like the precision). */
int  hdnnp_ctx_set_pair_record(hdnnp_ctx* c, int enable);
/* *active = 1 when symmetry-function evaluations of the current precision,
 * model and neighbour list record the pairs.  No counterpart in the
 * reference. */
int  hdnnp_ctx_pair_record_active(hdnnp_ctx* c, int* active);
/* Propagates a model change made after hdnnp_ctx_create (e.g.
 * hdnnp_model_set_scaling) to the context's device tables. */
int  hdnnp_ctx_update_model(hdnnp_ctx* c);

/* Tuning of the kernels' launch forms to the model and the hardware.  No
 * counterpart in the reference (its CPU kernels have none).  The knobs, as
 * four ints: SF members per group chunk of the device tables (1..16,
 * default 16), least waves per SIMD of the specialised SF and force
 * kernels (0 = the compiler's choice, -1 = built-in default, up to 8) and
 * 16-atom tiles per wave of the MFMA networks (1, 2 or 4).  None changes
 * what is computed; the member chunking fixes the order of the per-member
 * terms, so results may differ in the last bits between chunkings. */
#define HDNNP_TUNE_SWEEP 1 /* sweep even if a choice is cached */
/* Sweeps the knobs on the loaded structure (and its neighbour list, built
 * with the model's cutoff if there is none yet) at the current precision:
 * every member chunking (that changes the groups) with
 * six paired waves bounds, timing SF and force kernels of `reps`
 * evaluations (force_mode HDNNP_FORCES_*), then network tiles 1, 2, 4 at
 * the fastest kernels.  Applies the fastest and caches it for the process
 * per (model shape, precision, device); without HDNNP_TUNE_SWEEP a cached
 * choice is applied at once (*nrows = 0, no structure needed).  Rows of 8
 * doubles into table (up to cap rows; *nrows = rows swept): chunk, SF
 * waves, force waves, network tiles, SF ms, force ms, network ms, chosen
 * (1: the SF choice, 2: the force choice, 3: both).  Ends an MD run. */
int  hdnnp_ctx_autotune(hdnnp_ctx* c, int force_mode, int flags, int reps,
                        double* table, int cap, int* nrows);
/* The context's knobs (out4) / sets them (in4; -1 waves = default).  */
int  hdnnp_ctx_tuning(hdnnp_ctx* c, int* out4);
int  hdnnp_ctx_set_tuning(hdnnp_ctx* c, const int* in4);
/* Name of the SF (which = 0) or force (1) kernel the context launches at
 * its precision and tuning, as profilers list it: the specialised
 * hdnnp_plan_{sf,force}_{f32,f64}_m<chunk>_w<waves> (which compiles it if
 * needed) or the generic k_sf_forward / k_force_backward.  No counterpart
 * in the reference. */
int  hdnnp_ctx_kernel_name(hdnnp_ctx* c, int which, char* buf, int cap);

/* --------------------------------------------------------- structure -- */

/* Loads a structure: n atoms, positions r (n x 3, row-major), element
 * indices elem (n), box = three lattice vectors as rows (9 doubles) or NULL
 * for a non-periodic structure.  Positions must already lie in the cell as
 * after Structure::remap (Structure.cpp:1223-1247), which is what
 * Structure::readFromLines leaves (Structure.cpp:268).  Replaces the
 * Structure fields atoms[i].r, atoms[i].element, box, isPeriodic that Mode
 * reads (Structure.h:67-118). */
int hdnnp_set_structure(hdnnp_ctx* c, int64_t n, const double* r,
                        const int32_t* elem, const double* box);
/* Loads a batch of nstruct independent structures, evaluated together by
 * every following call.  Replaces the per-structure loop
 *   for (Structure& s : structures) mode.evaluateNNP(s, useForces, true);
 * of nnp-dataset.cpp:206-211 (Mode::evaluateNNP, Mode.cpp:1967-1990) with
 * one launch per kernel over all atoms.  Structure s holds the atoms
 * [atom_offsets[s], atom_offsets[s+1]) of the flat arrays r (n x 3) and
 * elem (n), n = atom_offsets[nstruct]; atom_offsets[0] = 0 and every
 * structure has at least one atom.  boxes (9 per structure, lattice vectors
 * as rows) may be NULL for a batch of open structures; periodic (nstruct,
 * may be NULL = all periodic when boxes is given) marks which structures
 * use their box.  Each structure gets its own neighbour list with its own
 * image counts, exactly as Structure::calculateNeighborList would build it.
 * hdnnp_energy / hdnnp_evaluate return the batch total; per-structure
 * energies come from hdnnp_get_structure_energies; forces, G, dE/dG and
 * the neighbour list are over the flat atom index. */
int hdnnp_set_structures(hdnnp_ctx* c, int64_t nstruct,
                         const int64_t* atom_offsets, const double* r,
                         const int32_t* elem, const double* boxes,
                         const int32_t* periodic);
/* Number of structures loaded (1 after hdnnp_set_structure). */
int64_t hdnnp_num_structures(const hdnnp_ctx* c);
/* Structure::energy of every structure (nstruct doubles) after the
 * networks; each is the sum of its atoms' energies in a fixed order. */
int hdnnp_get_structure_energies(hdnnp_ctx* c, double* energies);
/* Structure::pbc of every structure (3 x nstruct) of the last list. */
int hdnnp_get_structure_pbc(hdnnp_ctx* c, int32_t* pbc);
/* New positions for the same atoms (host array, n x 3). */
int hdnnp_set_positions(hdnnp_ctx* c, const double* r);
/* New positions already on the device (n x 3 doubles); no copy to host. */
int hdnnp_set_positions_device(hdnnp_ctx* c, const double* r_device);

/* ----------------------------------------------------------- compute -- */

/* Structure::calculateNeighborList(cutoff, false) (Structure.cpp:287-399):
 * per atom all (j, image) with |r_i - r_j + image| <= cutoff, ordered by j
 * then image components ascending; bit-identical index, d and dr.
 * Built with a cell list on the device. */
int hdnnp_neighbor_list(hdnnp_ctx* c, double cutoff);
/* Mode::calculateSymmetryFunctionGroups(structure, derivatives)
 * (Mode.cpp:1561-1640).  Computes the scaled SF values G.  With
 * derivatives = 1 the per-neighbour dG/dr of the reference are not stored:
 * the force kernel recomputes them (see hdnnp_forces). */
int hdnnp_symmetry_functions(hdnnp_ctx* c, int derivatives);
/* Mode::calculateAtomicNeuralNetworks(structure, derivatives)
 * (Mode.cpp:1642-1663): atomic energies and, with derivatives, dE_i/dG_i. */
int hdnnp_atomic_networks(hdnnp_ctx* c, int derivatives);
/* Mode::calculateEnergy (Mode.cpp:1803-1828): energy = sum of atomic
 * energies, summed in FP64 with a fixed order. */
int hdnnp_energy(hdnnp_ctx* c, double* energy);
/* Mode::calculateForces (Mode.cpp:1849-1913) with Atom::
 * calculateSelfForceShort / calculatePairForceShort (Atom.cpp:371-401):
 * F = -dE/dr by contracting dE/dG with dG/dr recomputed on the fly and a
 * deterministic per-atom gather (no atomics).  mode = HDNNP_FORCES_*. */
int hdnnp_forces(hdnnp_ctx* c, int mode);
/* Mode::evaluateNNP (Mode.cpp:1967-2016): neighbour list with the model's
 * maximum cutoff, SFs, networks, energy and (if force_mode >= 0) forces. */
int hdnnp_evaluate(hdnnp_ctx* c, int force_mode, double* energy);

/* ----------------------------------------------------------- results -- */

int64_t hdnnp_num_atoms(const hdnnp_ctx* c);
/* Total neighbour-list entries after hdnnp_neighbor_list. */
int64_t hdnnp_num_neighbor_entries(const hdnnp_ctx* c);
/* Structure::pbc (Structure.cpp:486-517) of the last neighbour list. */
int hdnnp_get_pbc(hdnnp_ctx* c, int32_t* pbc);
/* Neighbour list as CSR: offsets (n+1), index (M), image (3M, the integer
 * image vector bc of Structure.cpp:309-322), d (M), dr (3M).  Any output
 * pointer may be NULL.  Replaces Atom::neighbors / Atom::Neighbor
 * (Atom.h:37-79). */
int hdnnp_get_neighbors(hdnnp_ctx* c, int64_t* offsets, int32_t* index,
                        int32_t* image, double* d, double* dr);
/* Atom::G (n x ld, ld >= max SF count, zero padded). */
int hdnnp_get_G(hdnnp_ctx* c, double* G, int ld);
/* Atom::dEdG (n x ld). */
int hdnnp_get_dEdG(hdnnp_ctx* c, double* dEdG, int ld);
/* Atom::energy (n). */
int hdnnp_get_atomic_energies(hdnnp_ctx* c, double* e);
/* Atom::f (n x 3). */
int hdnnp_get_forces(hdnnp_ctx* c, double* f);
/* Per-atom self force -sum_s dE_i/dG_s dG_s/dr_i (n x 3) and per
 * neighbour-list entry pair force -sum_s dE_i/dG_s dG_s/dr_(slot) (M x 3),
 * i.e. Atom::calculateSelfForceShort and Atom::calculatePairForceShort for
 * every slot (Atom.cpp:371-401), after hdnnp_forces.  Either pointer may
 * be NULL.  These are what a caller scattering forces to ghost atoms needs
 * (InterfaceLammps::getForces, InterfaceLammps.cpp:546-581). */
int hdnnp_get_pair_forces(hdnnp_ctx* c, double* self_forces,
                          double* pair_forces);
/* ------------------------------------------- caller-given neighbours -- */

/* The InterfaceLammps path (src/libnnpif/LAMMPS/InterfaceLammps.cpp:267-621,
 * pair_hdnnp.cpp:81-114): the caller hands over, instead of positions and
 * a cell, ncentre centre atoms with their neighbour lists, as
 * InterfaceLammps::setLocalAtoms + addNeighbor do.  centre_index (ncentre,
 * NULL = 0..ncentre-1) and nbr_index (M) are the caller's atom indices
 * (local or ghost), used only to scatter forces; centre_elem (ncentre) and
 * nbr_elem (M) element indices; offsets (ncentre + 1, offsets[0] = 0,
 * M = offsets[ncentre]) delimit each centre's entries; dr (3M) holds
 * r_centre - r_neighbour in the model's internal units.  Any order of the
 * entries is accepted (the SF sums are then formed in that order).  The
 * list replaces the structure: hdnnp_symmetry_functions,
 * hdnnp_atomic_networks, hdnnp_energy, hdnnp_forces and hdnnp_evaluate
 * work on the centres; hdnnp_neighbor_list, hdnnp_set_positions and
 * hdnnp_get_forces are errors until hdnnp_set_structure is called. */
int hdnnp_set_neighbor_list(hdnnp_ctx* c, int64_t ncentre,
                            const int32_t* centre_index,
                            const int32_t* centre_elem,
                            const int64_t* offsets, const int32_t* nbr_index,
                            const int32_t* nbr_elem, const double* dr);
/* InterfaceLammps::getForces (InterfaceLammps.cpp:546-581) after
 * hdnnp_forces on a caller-given list: adds every centre's self force to
 * f[centre_index] and every entry's pair force to f[nbr_index] (f has nout
 * rows of 3, internal units), centres in order, entries in list order. */
int hdnnp_get_scattered_forces(hdnnp_ctx* c, int64_t nout, double* f);
/* No counterpart in the reference.  hdnnp_get_scattered_forces for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation): the same kernel and the
 * same result bit for bit, with the forces staged in a fixed 12 MiB device
 * array of the library instead of a buffer the context allocates.  At most
 * 524,288 caller atoms may receive forces (more is HDNNP_E_CAPACITY).  Every
 * call on one GPU shares that array, so calls that overlap on one GPU
 * race. */
int hdnnp_get_scattered_forces_noalloc(hdnnp_ctx* c, int64_t nout, double* f);

/* Device pointer of the n x 3 force array (valid until the next
 * hdnnp_set_structure); for callers that keep MD state on the GPU. */
const double* hdnnp_forces_device(hdnnp_ctx* c);
/* ---------------------------------------------------------------- MD -- */

/* Velocity-Verlet NVE on the device.  No counterpart in libnnp: replaces
 * the loop LAMMPS runs around pair_style hdnnp (fix nve with
 * pair_hdnnp.cpp:81-114 computing forces each step) and its neighbour
 * policy "neighbor <skin> bin; neigh_modify every 1 delay 0 check yes".
 * The list is built with rc + skin and rebuilt (atoms wrapped into the
 * cell) when an atom has moved more than skin / 2 since the last build;
 * every step filters it to the entries within rc, in the reference's
 * order, so each force evaluation sees the list a fresh build would give.
 * Each step is two HIP graphs (kick + drift + filter; forces + kick +
 * energy record) with a 2-int flag read in between; use_graph = 0 issues
 * the same kernels one by one.
 *
 * hdnnp_md_init: the loaded periodic structure, masses per element,
 * time step dt, skin, ftoa = acceleration per (force / mass) in the
 * caller's units (eV/A/amu with dt in fs: 9.648533212e-3), initial
 * velocities v (n x 3, NULL = zero), force mode (HDNNP_FORCES_*).  It
 * computes the forces at the initial positions.  hdnnp_set_structure,
 * hdnnp_set_positions*, hdnnp_neighbor_list, hdnnp_evaluate and
 * hdnnp_set_neighbor_list end the run. */
int hdnnp_md_init(hdnnp_ctx* c, const double* masses, double dt, double skin,
                  double ftoa, const double* v, int force_mode, int use_graph);
/* nsteps steps.  epot[k] (the model's energy) and ekin[k] = sum m v^2 / 2
 * / ftoa after step k, each may be NULL. */
int hdnnp_md_run(hdnnp_ctx* c, int64_t nsteps, double* epot, double* ekin);
/* Positions (unwrapped since the last rebuild) and velocities, n x 3. */
int hdnnp_md_get_state(hdnnp_ctx* c, double* r, double* v);
/* Neighbour-list builds since hdnnp_md_init (the first one included). */
int64_t hdnnp_md_rebuilds(const hdnnp_ctx* c);
/* Filtered entries whose skin-list entry had no partner since
 * hdnnp_md_init: 0 for a consistent list (such entries are left out of
 * the force gather). */
int64_t hdnnp_md_partner_misses(hdnnp_ctx* c);

/* Multiple-time-step NVE (r-RESPA, Tuckerman, Berne & Martyna 1992) with
 * a cheap surrogate model `sur` as the fast force and the full model
 * `full` as the reference: each outer step (dt_outer) is a half kick with
 * F_full - F_sur, n_inner velocity-Verlet substeps of dt_outer / n_inner on
 * the surrogate, a full evaluation and a second half kick.  After every
 * outer step the force RMSE of the surrogate against the full model is
 * measured; when it exceeds rmse_gate the next outer step runs the full model
 * for all n_inner substeps instead.  Both contexts hold the same periodic
 * structure; masses, skin, ftoa, v and use_graph as for hdnnp_md_init.
 * The state lives in `sur` (hdnnp_md_get_state(sur, ...)).  No counterpart
 * in libnnp. */
int hdnnp_mts_init(hdnnp_ctx* full, hdnnp_ctx* sur, const double* masses,
                   double dt_outer, int n_inner, double skin, double ftoa,
                   const double* v, double rmse_gate, int use_graph);
/* nouter outer steps; per step (each array may be NULL): the full model's
 * energy and the kinetic energy at the end of the step, the measured force
 * RMSE and the mode of the step (0 surrogate substeps, 1 full substeps). */
int hdnnp_mts_run(hdnnp_ctx* sur, int64_t nouter, double* epot, double* ekin,
                  double* rmse, int32_t* mode);

/* ---------------------------------------------------------- training -- */

/* Training on energies and forces (the distilled surrogate of the
 * multiple-time-step integrator, or any 2G model of the supported SFs).  No counterpart in the
 * 2G prediction path; the reference trains with libnnptrain
 * (Training.cpp), which this does not mirror.  Weights are the flat array
 * of every element's network in element order, per layer W[k][j]
 * row-major then b[k] (the order of the weight files is per layer j-major,
 * NeuralNetwork.cpp:147-169). */
int64_t hdnnp_num_weights(const hdnnp_ctx* c);
int hdnnp_get_weights(hdnnp_ctx* c, double* w);
int hdnnp_set_weights(hdnnp_ctx* c, const double* w);
/* Evaluates the loaded structure (FP64, consistent forces) and writes the
 * gradient of L = w_e ((E - e_ref) / N)^2 + w_f / (3N) sum |F - f_ref|^2
 * with respect to every weight (grad: hdnnp_num_weights doubles, may be
 * NULL; it is kept for hdnnp_train_adam).  loss (2 doubles, may be NULL)
 * = the two terms.  Every SF type; identity /
 * tanh / logistic / softplus layers with one output (HDNNP_E_UNSUPPORTED
 * otherwise). */
int hdnnp_train_gradient(hdnnp_ctx* c, const double* f_ref, double e_ref,
                         double w_e, double w_f, double* grad, double* loss);
/* One Adam step (Kingma & Ba 2015) with the last gradient. */
int hdnnp_train_adam(hdnnp_ctx* c, double lr, double beta1, double beta2,
                     double eps);
/* Restart Adam (zero moments at the next step). */
int hdnnp_train_reset(hdnnp_ctx* c);
/* Kalman filter on the weights (KalmanFilter.cpp, standard type): the
 * error covariance P = I / epsilon (hdnnp_num_weights squared doubles on
 * the device), process noise q0 annealed by exp(-qtau) per update down to
 * qmin, learning rate eta0 raised by exp(etatau) per update up to etamax
 * (KalmanFilter::setParametersStandard, KalmanFilter.cpp:199-220). */
int hdnnp_kalman_init(hdnnp_ctx* c, double epsilon, double q0, double qtau,
                      double qmin, double eta0, double etatau, double etamax);
/* One update with a scalar observation (KalmanFilter::update,
 * KalmanFilter.cpp:126-196): Jacobian h (hdnnp_num_weights doubles, the
 * flat weight order, host memory) and innovation xi = observed - predicted.
 * X = P h, K = X / (h.X + 1/eta), P -= K X^T, P += q I, w += K xi. */
int hdnnp_kalman_update(hdnnp_ctx* c, const double* h, double xi);
/* Derivative of every structure's energy with respect to every network
 * weight, for any model (only the SFs enter): J[s * nw + w], nw =
 * hdnnp_num_weights, s over the loaded structures (one, or the batch of
 * hdnnp_set_structures).  Replaces, per structure, the sum over its atoms
 * of NeuralNetwork::calculateDEdc (NeuralNetwork.cpp:444-490) as training
 * forms it (Training.cpp, energy update).  file_order != 0: weights in the
 * reference's connection order (NeuralNetwork.cpp:147-169, element after
 * element); 0: the engine's flat order of hdnnp_get_weights.  Needs the
 * neighbour list; normalize_nodes is not supported. */
int hdnnp_energy_weight_jacobian(hdnnp_ctx* c, double* J, int file_order);

/* Number of extrapolation warnings of the last SF evaluation: G outside
 * [Gmin, Gmax] +- 1000 eps before scaling (Element.cpp:462-541).
 * per_atom (n, may be NULL) receives the count per atom. */
int64_t hdnnp_extrapolation_warnings(hdnnp_ctx* c, int32_t* per_atom);
/* The warnings themselves, one record per (atom, SF) in the order
 * Element::updateSymmetryFunctionStatistics meets them (atoms in order, SFs
 * in NN-input order; Element.cpp:479-539 called from Mode.cpp:1625-1633):
 * atom index, SF index within the atom's element (SymFnc::getIndex), the
 * unscaled value (SymFnc::unscale) and the SF's Gmin and Gmax.  Up to max
 * records are written (any array may be NULL); returns the total number. */
int64_t hdnnp_extrapolation_warning_list(hdnnp_ctx* c, int64_t max,
                                         int32_t* atom, int32_t* sf,
                                         double* value, double* gmin,
                                         double* gmax);

/* Symmetry-function statistics over the loaded structure as nnp-scaling
 * writes them to scaling.data (Dataset.cpp:974-1100): SFs evaluated without
 * scaling, per SF of element e: min, max, mean, sigma (4 doubles each).
 * Needs hdnnp_neighbor_list first.  Returns the number of atoms of e. */
int64_t hdnnp_sf_statistics(hdnnp_ctx* c, int e, double* out);
/* The accumulators behind it (SymFncStatistics::Container,
 * SymFncStatistics.h:28-147): per SF min, max, sum, sum of squares, for
 * combining several loads (batches of structures) into one statistics set,
 * as nnp-scaling combines structures and MPI ranks
 * (Dataset::collectSymmetryFunctionStatistics).  Returns the atom count. */
int64_t hdnnp_sf_statistics_raw(hdnnp_ctx* c, int e, double* out);

/* ---------------------------------------------------------- profiling -- */

/* Average device time in milliseconds of each kernel class over the calls
 * since the last reset, measured with HIP events on the context's stream.
 * names: "neighbor", "symfunc", "network", "force", "gather" and "partner"
 * (the partner map of a list build, which runs on the context's second
 * stream while the SF, network and force kernels run).  Returns -1 for an
 * unknown name.  enable != 0 turns event recording on. */
int    hdnnp_profile_enable(hdnnp_ctx* c, int enable);
double hdnnp_profile_avg_ms(hdnnp_ctx* c, const char* name, int64_t* calls);
int    hdnnp_profile_reset(hdnnp_ctx* c);
/* Algorithmic work of the loaded structure and neighbour list, counted on
 * the device with the filters of SymGrpExpRad/Angn/Angw::calculate, for
 * roofline accounting (HDNNP_WORK_COUNTS values):
 *   [0] neighbour-list entries          [1] radial member evaluations
 *   [2] narrow angular pairs tested     [3] wide angular pairs tested
 *       (element-matched, rij, rik < rc)
 *   [4] narrow triplets (rjk < rc)      [5] wide triplets
 *   [6] narrow member evaluations       [7] wide member evaluations
 *   [8] multiplies of pow_int(1 + lambda cos, zeta - 1) over [6] + [7]
 *   [9] dE/dG . dG/dr terms of Mode::calculateForces (per atom: SF count +
 *       per neighbour slot the SF-derivative count of its element)
 * Needs hdnnp_neighbor_list first. */
#define HDNNP_WORK_COUNTS 10
int    hdnnp_work_counts(hdnnp_ctx* c, int64_t* counts);

#ifdef __cplusplus
}
#endif
#endif /* HDNNP_H */
