// Model-specialised kernels compiled at run time (see specialize.hip).
#ifndef HDNNP_SPECIALIZE_H
#define HDNNP_SPECIALIZE_H

#include <hip/hip_runtime.h>

#include <string>

#include "model.h"
#include "tables.h"

namespace hdnnp {

// Kernels taking SFArgs / ForceArgs by value, launched with 64 threads per
// atom and the same dynamic LDS as the generic kernels.  Null if the
// specialisation is unavailable (then the generic kernels are used).
struct PlanKernels
{
    hipFunction_t sf = nullptr;
    hipFunction_t force = nullptr;
};

// Occupancy bounds of the specialised kernels: the least waves per SIMD
// the compiler must leave room for (amdgpu_waves_per_eu), 0 = its own
// choice, -1 = the built-in defaults.  The autotuner's knobs, with the
// group chunk size of the model tables.
struct PlanTuning
{
    int sfWaves = -1;
    int forceWaves = -1;
};

// Specialised kernels for the model (fp32: mixed-precision arithmetic) on
// the given device, compiled on first request and kept for the process.
PlanKernels plan_kernels(Model::Flat const& F, ModelTables const& T, int fp32,
                         int device, PlanTuning const& tune = PlanTuning());

// Name of the specialised SF (force = false) or force kernel for the
// tables, arithmetic and tuning: hdnnp_plan_{sf,force}_{f32,f64}_m<members
// per chunk>_w<waves-per-SIMD bound, 0 = compiler's choice>.
std::string plan_kernel_name(ModelTables const& T, int fp32,
                             PlanTuning const& tune, bool force);

// The generated member-loop code (the part of the kernel source that
// depends on the model).
std::string plan_source(Model::Flat const& F);

// False when the environment sets N2P2_AMD_RTC=0.
bool plan_enabled_by_default();
// False when the environment sets N2P2_AMD_PAIR_RECORD=0.
bool pair_record_enabled_by_default();

}  // namespace hdnnp

#endif
