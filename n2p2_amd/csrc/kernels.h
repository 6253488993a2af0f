// Host-callable launchers of the HIP kernels (one translation unit each).
#ifndef HDNNP_KERNELS_H
#define HDNNP_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernel_args.h"
#include "tables.h"

namespace hdnnp {

// Cell-list geometry for the neighbour build.
struct NLParams
{
    int    n;
    int    periodic;
    double box[3][3];      // lattice vectors as rows
    double invbox[3][3];   // Structure::invbox (Structure.cpp:519-545)
    double origin[3];      // non-periodic: lower corner of the bounding box
    int    nc[3];          // cells per direction
    int    m[3];           // cells searched on each side
    double rc2;
};

hipError_t scan_exclusive(const int* in, int* out, int n, int* blockSums,
                          int* total, hipStream_t s);
// out[0] = maximum of in[0..n).
hipError_t launch_max_int(const int* in, int n, int* out, hipStream_t s);
hipError_t launch_nl_bin(NLParams const& p, const double* r, int* cellOfAtom,
                         int* rank, int* cellCount, hipStream_t s);
hipError_t launch_nl_scatter(NLParams const& p, const int* cellOfAtom,
                             const int* rank, const int* cellStart,
                             int* cellAtoms, hipStream_t s);
hipError_t launch_nl_count(NLParams const& p, const double* r, const int* elem,
                           const int* cellOfAtom, const int* cellStart,
                           const int* cellCount, const int* cellAtoms,
                           int* count, int* stats, hipStream_t s);
hipError_t launch_nl_fill(NLParams const& p, const double* r, const int* elem,
                          const int* cellOfAtom, const int* cellStart,
                          const int* cellCount, const int* cellAtoms,
                          int* stats, const int* off, int keyCap,
                          uint64_t* keys, int* nbrJ, double4* geo,
                          hipStream_t s);
// Cell-per-workgroup variants (at most 64 neighbour cells searched):
// stats[2] receives the largest staged candidate count when a cell exceeds
// the kernel's LDS capacity (the caller then uses the per-atom kernels).
hipError_t launch_nl_count_cell(NLParams const& p, int ncell, const double* r,
                                const int* cellStart, const int* cellCount,
                                const int* cellAtoms, int* count, int* stats,
                                hipStream_t s);
// Writes atom i's entries from off[i] on, at most limit of them; count[i]
// receives the true count, stats[3] = 1 if a count exceeds limit (the
// largest count: launch_max_int over count).
hipError_t launch_nl_fill_cell(NLParams const& p, int ncell, const double* r,
                               const int* cellStart, const int* cellCount,
                               const int* cellAtoms, int* count, int* stats,
                               const int* off, int limit, uint64_t* keys,
                               int* nbrJ, double4* geo, hipStream_t s);
// Cell of one structure of a batch.  The host sets everything but the
// atoms' fractional extent fmin/fmax, which launch_batch_extent fills.
struct BatchCell
{
    double box[3][3];      // lattice vectors as rows (zero if open)
    double invbox[3][3];   // its inverse: fractional f = r invbox
    double proj[3];        // distance between opposite faces
    double fmin[3], fmax[3];
    int    pbc[3];         // images searched per direction
    int    periodic;
};
hipError_t launch_batch_extent(int nstruct, const double* r,
                               const int* structOff, BatchCell* cells,
                               hipStream_t s);
// Batched structures: per-atom lists within each atom's structure
// (structOf, structOff) over its cell's images.  mode 0 counts into count;
// mode 1 writes at most `limit` entries per atom from off, stores count
// and sets stats[3] on overflow (the largest count: launch_max_int).
hipError_t launch_nl_batch(int mode, int n, double rc, const double* r,
                           const int* structOf, const int* structOff,
                           const BatchCell* cells, int* count, int* stats,
                           const int* off, int limit, uint64_t* keys,
                           int* nbrJ, double4* geo, hipStream_t s);
// --- Training of a radial-only model (kernels_train.hip)
struct TrainArgs
{
    ModelTables T;
    int n;
    const int* off;
    const int* cnt;
    const int* nbrJ;
    const double4* geo;
    const int* elem;
    const double* G;        // scaled SFs of the last evaluation
    int ldG;
    const double* F;        // forces of the last evaluation
    const double* Fref;     // reference forces
    double* U;              // n x ldG: u_as
    double* d2;             // n: |F - Fref|^2
    const double* energy;   // device: E of the last evaluation
    double Eref, wE, cF;    // cF = -2 wF / (3N)
};
int train_atoms_per_block();
int train_atoms_per_grad_block(ElemDesc const& ed);
size_t train_lds_per_atom(ElemDesc const& ed);
hipError_t launch_train_u(TrainArgs const& A, hipStream_t s);
// The compact SF types' part of u (after launch_train_u); cap = the largest
// neighbour count.
hipError_t launch_train_u_compact(TrainArgs const& A, int cap, hipStream_t s);
size_t train_u_compact_lds_bytes(ModelTables const& T, int cap);
hipError_t launch_train_grad(TrainArgs const& A, ElemDesc const& ed, int e,
                             const int* atoms, int na, double* part, int nw,
                             double* grad, hipStream_t s);
hipError_t launch_adam(int n, double* w, float* w32, const double* g,
                       double* m, double* v, double lr, double b1, double b2,
                       double eps, int64_t step, hipStream_t s);
hipError_t launch_weights_to_f32(int n, const double* w, float* w32,
                                 hipStream_t s);
// Kalman filter on the flat weights, one scalar observation: P (nw x nw)
// = I diag; update with Jacobian h (device), innovation xi, measurement
// noise rinv = 1/eta and process noise q (KalmanFilter.cpp:126-196).
hipError_t launch_kf_init(int nw, double* P, double diag, hipStream_t s);
hipError_t launch_kf_update(int nw, double* P, const double* h, double* X,
                            double* a, double rinv, double q, double xi,
                            double* w, float* w32, hipStream_t s);
// dE_s/dw per structure (blocks of at most train_atoms_per_block() atoms of
// one element and one structure; LDS perAtom doubles per atom).
hipError_t launch_energy_wgrad(ModelTables const& T, const double* G, int ldG,
                               const int* blkAtoms, const int* blk0,
                               const int* blkN, const int* blkElem, int nblk,
                               int perAtom, double* part, int maxP,
                               const int* grp0, const int* base,
                               const int* nwe, int ne, int nstruct,
                               int nwTotal, double* J, hipStream_t s);

// --- MD with a Verlet-skin list (kernels_md.hip)
struct MDCell
{
    double box[3][3];      // lattice vectors as rows
    double invbox[3][3];
};
hipError_t launch_md_kick(int n, double* v, const double* F, const int* elem,
                          const double* accel, double h, hipStream_t s);
hipError_t launch_md_drift(int n, double* r, const double* v, const double* r0,
                           double dt, double half2, int* flags, hipStream_t s);
hipError_t launch_md_clear_flags(int* flags, hipStream_t s);
hipError_t launch_md_kick_diff(int n, double* v, const double* Ff,
                               const double* Fs, const int* elem,
                               const double* accel, double h, hipStream_t s);
hipError_t launch_md_fdiff2(int n, const double* F1, const double* F2,
                            double* d2, hipStream_t s);
// moved[0] is set to 1 if an atom was shifted (the caller clears it).
hipError_t launch_md_wrap(int n, double* r, MDCell const& C, int* moved,
                          hipStream_t s);
hipError_t launch_md_filter(int n, MDCell const& C, double rc2, const double* r,
                            const int* off, const int* skinCnt,
                            const uint64_t* skinKeys, int* cnt, int* nbrJ,
                            double4* geo, int* newIdx, int cap, int* flags,
                            hipStream_t s);
// keys[newIdx[s]] = skinKeys[s] for the kept skin slots.
hipError_t launch_md_keys(int n, const int* off, const int* skinCnt,
                          const uint64_t* skinKeys, const int* newIdx,
                          uint64_t* keys, hipStream_t s);
// Force gather over a filtered list through the skin list's partners.
hipError_t launch_gather_md(int n, const int* off, const int* skinCnt,
                            const int* newIdx, const int* skinPartner,
                            const uint64_t* skinKeys, const double* cslot,
                            const double* cself, const int* order, int mode,
                            int64_t nslots, double* F, int* misses,
                            hipStream_t s);
int md_mv2_blocks(int n);
hipError_t launch_md_record(int n, const double* v, const int* elem,
                            const double* mass, double* part,
                            const double* energy, double ekinScale,
                            int* counter, int capRec, double* epot,
                            double* ekin, hipStream_t s);
// Per-structure energy sums (one block per structure, fixed order).
hipError_t launch_energy_segments(const double* E, const int* structOff,
                                  int nstruct, double* out, hipStream_t s);
// off[i] = i * slab for i <= n.
hipError_t launch_slab_starts(int n, int slab, int* off, hipStream_t s);
// Entries of atom i: [off[i], off[i] + cnt[i]).  order: the atoms sorted
// by cell (processing order only).
// Partner slots of a symmetric list (nslots = slots allocated; all are
// set to -1 first).
hipError_t launch_nl_partner(int n, const int* off, const int* cnt,
                             const uint64_t* keys, const int* order,
                             int* partner, int64_t nslots, hipStream_t s);

hipError_t launch_sf_forward(SFArgs const& a, hipStream_t s);
size_t sf_forward_lds_bytes(int cap, int fp32);
// Work counters (10 x uint64, accumulated): see k_sf_count.
hipError_t launch_sf_count(SFArgs const& a, unsigned long long* counters,
                           hipStream_t s);

hipError_t launch_force_backward(ForceArgs const& a, hipStream_t s);

// Compact symmetry functions (types 20-25), FP64: values after
// launch_sf_forward, per-slot force contributions added to cslot / cself
// after launch_force_backward.
hipError_t launch_sf_compact(SFArgs const& a, hipStream_t s);
hipError_t launch_force_compact(ForceArgs const& a, hipStream_t s);
size_t sf_compact_lds_bytes(ModelTables const& T, int cap);
size_t force_compact_lds_bytes(ModelTables const& T, int cap);
size_t force_backward_lds_bytes(int cap, int fp32);

// Entries of atom i: [off[i], off[i] + cnt[i]).  order: the atoms sorted
// by cell (processing order only).
// Forces of a caller-given list onto the caller's atoms, in the order of
// InterfaceLammps::getForces (k_scatter_ext).
hipError_t launch_scatter_ext(int nt, const int64_t* srcOff, const int* src,
                              const double* cslot, const double* cself,
                              double* f, hipStream_t s);
hipError_t launch_force_gather(int n, const int* off, const int* cnt,
                               const int* nbrJ, const int* partner,
                               const double* cslot, const double* cself,
                               const int* order, int mode, int64_t nslots,
                               double* F, hipStream_t s);

// Atomic networks for the atoms of one element.  mode -1: per-atom FP64
// kernel (any width); 0: FP64 MFMA tiles; 1: FP32 MFMA tiles; 2 / 3: FP16 /
// BF16 MFMA tiles (hi + lo operand split) with FP32 accumulation.  The MFMA
// kernels run `tiles` (1, 2 or 4) tiles of 16 atoms per wave and need
// nn_mfma_lds_bytes(ed, mode, tiles) of LDS.
struct NNArgs
{
    ModelTables T;
    ElemDesc ed;                 // host copy of T.elem[e]
    int mode;
    int e;                       // element
    const int* atoms;            // atom indices of element e
    int na;
    const double* G;
    int ldG;
    double* E;                   // per atom energy (indexed by atom)
    double* dEdG;                // n x ldG (indexed by atom)
    double* scratch;             // sum(width[l]) * 3 * na doubles
    int deriv;
    int tiles;                   // 16-atom tiles per wave (MFMA modes)
};
hipError_t launch_nn(NNArgs const& a, hipStream_t s);
size_t nn_mfma_lds_bytes(ElemDesc const& ed, int mode, int tiles);

hipError_t launch_energy_sum(const double* E, int n, double* out, hipStream_t s);

// Unscaled-SF statistics for one element (min, max, sum, sum of squares)
// over atoms of that element: out 4 x nsf.
hipError_t launch_sf_stats(const double* G, int ldG, const int* atoms, int na,
                           int nsf, double* out, hipStream_t s);

// Extrapolation-warning count per atom (Element.cpp:462-541 test).
hipError_t launch_ew(ModelTables T, const double* G, int ldG, const int* elem,
                     int n, int* perAtom, hipStream_t s);

}  // namespace hdnnp

#endif
