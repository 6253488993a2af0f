/*
 * Oracle: CPU restatement of n2p2's 2G-HDNNP energy-and-force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine
 * in n2p2_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to compute expected values or to
 * time the CPU baseline; the product never links or calls it.
 *
 * Every function follows the reference loop-for-loop (citations are given in
 * oracle.cpp, relative to the n2p2 source tree).  The C ABI below is what the
 * Python tests bind with ctypes.
 */
#ifndef HDNNP_ORACLE_H
#define HDNNP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model  orc_model;
typedef struct orc_struct orc_struct;

/* Last error message of the calling thread ("" if none). */
const char* orc_last_error(void);

/* Model = settings + scaling + weights, as Mode::setupGeneric +
 * setupSymmetryFunctionScaling + setupNeuralNetworkWeights.
 * scaling_file may be NULL/"" for "no scaling" (setupSymmetryFunctionScalingNone).
 * weights_fmt is a printf format with one %zu / %03zu for the atomic number. */
orc_model* orc_model_load(const char* settings_file, const char* scaling_file,
                          const char* weights_fmt);
void       orc_model_free(orc_model* m);
int        orc_model_num_elements(const orc_model* m);
/* Atomic number of element index e (elements are sorted by Z). */
int        orc_model_atomic_number(const orc_model* m, int e);
int        orc_model_num_sf(const orc_model* m, int e);
double     orc_model_max_cutoff(const orc_model* m);
int        orc_model_normalize(const orc_model* m, double* mean_energy,
                               double* conv_energy, double* conv_length);
/* Sorted SF table of element e: per SF 12 doubles:
 * type, e1, e2, eta, rs, lambda, zeta, rc, cutoff_type, cutoff_alpha,
 * scaling_factor, line_number.  out must hold 12*num_sf doubles. */
int        orc_model_sf_table(const orc_model* m, int e, double* out);
/* Group layout of element e: returns the number of groups; if out != NULL,
 * writes for each group (type, e1, e2, num_members, member SF indices...)
 * as int32 into out (capacity cap). */
int        orc_model_groups(const orc_model* m, int e, int32_t* out, int cap);

/* Structure from positions (n x 3), element indices, and box (3 x 3 rows =
 * lattice vectors) or NULL for non-periodic.  Positions are used as given
 * (apply orc_struct_remap for the reference's read-time wrapping). */
orc_struct* orc_struct_create(const orc_model* m, int64_t n, const double* r,
                              const int32_t* elem, const double* box);
/* Structure::readFromFile of the first configuration (includes remap). */
orc_struct* orc_struct_read(const orc_model* m, const char* path);
void        orc_struct_free(orc_struct* s);
int64_t     orc_struct_num_atoms(const orc_struct* s);
void        orc_struct_remap(orc_struct* s);
void        orc_struct_get_positions(const orc_struct* s, double* r);
void        orc_struct_get_elements(const orc_struct* s, int32_t* e);
void        orc_struct_get_box(const orc_struct* s, double* box);
int         orc_struct_to_normalized(orc_struct* s, const orc_model* m);

/* Structure::calculateNeighborList (brute force, cells=0) or the same list
 * built with a cell list (cells=1).  Returns total number of entries. */
int64_t orc_neighbor_list(orc_struct* s, double cutoff, int cells);
void    orc_get_pbc(const orc_struct* s, int32_t* pbc);
/* CSR export: offsets[n+1], index[M], d[M], dr[3M]. */
void    orc_get_neighbors(const orc_struct* s, int64_t* offsets, int32_t* idx,
                          double* d, double* dr);

/* Mode::calculateSymmetryFunctionGroups(structure, derivatives). */
int orc_symmetry_functions(orc_struct* s, const orc_model* m, int derivatives);
/* Mode::calculateAtomicNeuralNetworks(structure, derivatives). */
int orc_networks(orc_struct* s, orc_model* m, int derivatives);
/* Mode::calculateEnergy. */
double orc_energy(orc_struct* s);
/* Mode::calculateForces; mode 0 = reference gather (skips an atom's own
 * periodic self-image slots), mode 1 = adds those slots (consistent with
 * finite differences when the cell is smaller than the cutoff). */
int orc_forces(orc_struct* s, const orc_model* m, int mode);

/* Per-atom results.  G and dEdG are written as n x max_nsf (zero padded). */
int  orc_get_G(const orc_struct* s, double* G, int max_nsf);
int  orc_get_dEdG(const orc_struct* s, double* dEdG, int max_nsf);
void orc_get_atomic_energies(const orc_struct* s, double* e);
void orc_get_forces(const orc_struct* s, double* f);
/* dG/dr of atom i's SFs with respect to its own position: nsf x 3. */
int  orc_get_dGdr_self(const orc_struct* s, int64_t i, double* out);
/* Per-neighbour derivative vectors of atom i, slot k: numSFD x 3. Returns
 * numSFD. */
int  orc_get_dGdr_neighbor(const orc_struct* s, int64_t i, int64_t k,
                           double* out);

/* Atom::calculateSelfForceShort per atom (n x 3) and
 * Atom::calculatePairForceShort of every neighbour slot of every atom
 * (M x 3, CSR order), after orc_symmetry_functions(derivatives=1) and
 * orc_networks(derivatives=1). */
int orc_get_pair_forces(const orc_struct* s, const orc_model* m,
                        double* self_forces, double* pair_forces);

/* Prediction::readStructureFromFile + predict in one call: full pipeline in
 * physical units with energy offsets.  Writes energy and forces (n x 3). */
int orc_predict_file(orc_model* m, const char* path, double* energy,
                     double* forces, int64_t cap_atoms, int64_t* n_atoms);

/* Full evaluation (neighbour list, SFs, NNs, energy, forces) of a structure
 * in the model's internal units; returns energy. force_mode as orc_forces. */
double orc_evaluate(orc_struct* s, orc_model* m, int cells, int force_mode);
/* Energy only (no derivatives) of a structure; used by finite differences. */
double orc_energy_only(orc_struct* s, orc_model* m, int cells);

/* SF statistics over a structure with the model's SFs evaluated unscaled
 * (the nnp-scaling definition): per SF min, max, mean, sigma.
 * out holds 4*num_sf doubles for element e. */
int orc_sf_statistics(orc_struct* s, orc_model* m, int e, double* out);

/* Element::updateSymmetryFunctionStatistics over all atoms after
 * orc_symmetry_functions: extrapolation-warning records (atom, SF index,
 * unscaled value), up to cap written; returns the total. */
int64_t orc_extrapolation_warnings(const orc_struct* s, const orc_model* m,
                                   int32_t* atom, int32_t* sf, double* value,
                                   int64_t cap);

/* Central finite-difference forces for atoms in list[0..nlist): out 3*nlist.
 * Follows Dataset::prepareNumericForces + nnp-checkf (remap after shift). */
int orc_numeric_forces(orc_struct* s, orc_model* m, double delta,
                       const int64_t* list, int64_t nlist, int cells,
                       double* out);

/* Number of OpenMP threads the oracle uses. */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
