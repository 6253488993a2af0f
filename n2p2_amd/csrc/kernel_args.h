// Argument blocks of the symmetry-function and force kernels, shared by the
// ahead-of-time kernels and the model-specialised ones compiled at run time.
#ifndef HDNNP_KERNEL_ARGS_H
#define HDNNP_KERNEL_ARGS_H

#include "tables.h"

namespace hdnnp {

// Symmetry functions.  raw != 0 skips scaling (statistics); fp32 != 0
// selects the mixed-precision kernel (FP32 geometry and member sums).
struct SFArgs
{
    ModelTables T;
    int n;                       // atoms
    const int* off;              // first entry of each atom
    const int* cnt;              // entries of each atom
    const double4* geo;          // dx, dy, dz, d per entry
    const int* nbrJ;             // neighbour atom per entry
    const int* elem;             // element per atom
    double* G;                   // n x ldG
    int ldG;
    int cap;                     // LDS neighbour capacity per wave
    int raw;
    int fp32;
    // Pair record (null: none): per atom pcap entries, the first
    // kRecordGroups the passing-pair count of each angular group chunk
    // (kRecordMissing: not recorded), then the pairs in evaluation order as
    // staged indices j | k << 8.
    unsigned short* pairs;
    int pcap;
};

struct ForceArgs
{
    ModelTables T;
    int n;
    const int* off;              // first entry of each atom
    const int* cnt;              // entries of each atom
    const double4* geo;
    const int* nbrJ;
    const int* elem;
    const double* dEdG;          // n x ldG
    int ldG;
    double* cslot;               // 3 per entry (x, y, z), entry-indexed
    double* cself;               // 3 x n
    int cap;
    int fp32;                    // mixed-precision contraction
    // The SF kernel's pair record of the same list (null: walk the pairs).
    const unsigned short* pairs;
    int pcap;
};

}  // namespace hdnnp

#endif
