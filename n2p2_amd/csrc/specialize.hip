// Model-specialised symmetry-function and force kernels.
//
// The member loops of the two heavy kernels are generic: they read each
// member's lambda, zeta and exponential parameters from the group tables
// and branch on them per triplet.  For a given model these are constants,
// so this file writes them out as straight-line C++ (one switch case per
// angular group chunk: one exponential per (eta, rs) run, one chain of
// powers plambda^(2^l - 1) per (eta, rs, lambda) run, every member's
// contribution with its lambda zeta folded in), compiles the kernels
// together with that code through hipRTC for the context's device, and
// keeps the loaded module for the process.  The kernel bodies are the same
// device code as the ahead-of-time kernels (force_core.h, sf_core.h).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "model.h"
#include "specialize.h"

namespace hdnnp {

namespace {

// Sources handed to hipRTC as named headers (generated at build time).
#include "rtc_sources.inc"

std::string lit(double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "R(%a)", v);
    return b;
}

// plambda = 1 + lambda cos.
std::string plambda(double lambda)
{
    if (lambda == 1.0) return "R(1) + costijk";
    if (lambda == -1.0) return "R(1) - costijk";
    return "R(1) + " + lit(lambda) + " * costijk";
}

// Exponent sum of an (eta, rs) run into s_.
void emit_exp_sum(std::ostringstream& o, LamRun const& r, bool wide)
{
    if (r.rs > 0.0)
    {
        o << "        R const a_ = rij - " << lit(r.rs) << ";\n"
          << "        R const b_ = rik - " << lit(r.rs) << ";\n"
          << "        R s_ = a_ * a_ + b_ * b_;\n";
        if (!wide)
            o << "        R const c_ = rjk - " << lit(r.rs) << ";\n"
              << "        s_ += c_ * c_;\n";
    }
    else o << "        R const s_ = r2sum;\n";
}

// The member sums of one group chunk.  Every sum starts with an assignment
// of its first term, not "0 +" it: the compiler may not drop an addition of
// +0 (it would change -0 into +0), so a zero start costs an instruction per
// sum and triplet.
void emit_backward_case(std::ostringstream& o, int gi, GroupDesc const& g,
                        Model::Flat const& F)
{
    bool const wide = (g.type == 9);
    o << "    case " << gi << ":\n    {\n"
      << "        R S1, U, vexp;\n"
      << "        R E0 = R(0), Ejx = R(0), Ekx = R(0), E3x = R(0);\n";
    bool anyZero = false, anyShift = false;
    bool firstT = true, firstE0 = true, firstEx = true, firstS = true;
    double eta = 0.0, rs = 0.0;
    auto acc = [&](bool& first) { char const* op = first ? " = " : " += ";
                                  first = false; return op; };
    auto close_run = [&]() {
        bool ft = firstT;
        o << "        T1" << acc(ft) << "vexp * S1;\n"
          << "        { R const vu = vexp * U; V" << acc(firstT) << "vu;";
        if (rs > 0.0)
        {
            bool fx = firstEx, fy = firstEx;
            o << " R const te = " << lit(2.0 * eta) << " * vu;"
              << " Ejx" << acc(fx) << "te * a_ * invj; Ekx" << acc(fy)
              << "te * b_ * invk;";
            if (!wide)
            {
                bool fz = firstEx;
                o << " E3x" << acc(fz) << "te * c_ * invjk;";
            }
            firstEx = false;
        }
        else o << " E0" << acc(firstE0) << lit(2.0 * eta) << " * vu;";
        o << " }\n        }\n";
        firstS = true;
    };
    bool open = false;
    for (int ri = 0; ri < g.nrun; ++ri)
    {
        LamRun const& r = F.run[g.run0 + ri];
        if (r.newexp)
        {
            if (open) close_run();
            open = true;
            eta = r.eta;
            rs = r.rs;
            (rs > 0.0 ? anyShift : anyZero) = true;
            // The run's block stays open until it is closed, so that the
            // shifted distances a_, b_, c_ remain in scope.
            o << "        {\n";
            emit_exp_sum(o, r, wide);
            o << "        vexp = exp_scaled(exp_factor(" << lit(r.eta)
              << ") * s_);\n";
        }
        // Per lambda run: W_ = sum w P, Z_ = sum (w zeta lambda) P, two
        // FMAs per member; U += plambda W_ and S1 += Z_ at the run's end.
        o << "        {\n        R const pl = " << plambda(r.lambda) << ";\n"
          << "        R P = pl > R(0) ? R(1) : R(0), X = pl;\n"
          << "        R W_, Z_;\n";
        bool firstW = true;
        int level = 0;
        int m = r.m0;
        for (int l = 0; l <= r.lmax; ++l)
            for (; m < r.lvEnd[l]; ++m)
            {
                MemberDesc const& md = F.member[g.mem0 + m];
                for (; level < md.level; ++level)
                    o << "        P *= X; X *= X;\n";
                bool fz = firstW;
                o << "        W_" << acc(firstW) << "w[" << m << "] * P; Z_"
                  << acc(fz) << "wz[" << m << "] * P;\n";
            }
        for (; m < r.mend; ++m)
        {
            MemberDesc const& md = F.member[g.mem0 + m];
            o << "        { R const pw_ = ";
            if (md.useint) o << "pow_int(pl, " << md.zint - 1 << ")";
            else o << "pow(pl, " << lit(md.zeta - 1.0) << ")";
            bool fz = firstW;
            o << "; R const f_ = pl > R(0) ? pw_ : R(0); W_" << acc(firstW)
              << "w[" << m << "] * f_; Z_" << acc(fz) << "wz[" << m
              << "] * f_; }\n";
        }
        if (firstW) o << "        W_ = Z_ = R(0);\n";   // no members
        bool fu = firstS;
        o << "        S1" << acc(firstS) << "Z_; U" << acc(fu)
          << "W_ * pl;\n        }\n";
    }
    if (open) close_run();
    else o << "        T1 = V = R(0);\n";
    if (anyZero && !anyShift)
        o << "        Ej = E0; Ek = E0; E3 = E0;\n";
    else if (!anyZero)
        o << "        Ej = Ejx; Ek = Ekx; E3 = E3x;\n";
    else
        o << "        Ej = E0 + Ejx; Ek = E0 + Ekx; E3 = E0 + E3x;\n";
    o << "        (void)E0; (void)Ejx; (void)Ekx; (void)E3x;\n"
      << "        break;\n    }\n";
}

void emit_forward_case(std::ostringstream& o, int gi, GroupDesc const& g,
                       Model::Flat const& F)
{
    bool const wide = (g.type == 9);
    o << "    case " << gi << ":\n    {\n        R vp = R(0);\n";
    for (int ri = 0; ri < g.nrun; ++ri)
    {
        LamRun const& r = F.run[g.run0 + ri];
        if (r.newexp)
        {
            o << "        {\n";
            emit_exp_sum(o, r, wide);
            o << "        vp = exp_scaled(exp_factor(" << lit(r.eta)
              << ") * s_) * pfc;\n        }\n";
        }
        o << "        {\n        R const pl = " << plambda(r.lambda) << ";\n"
          << "        R P = pl > R(0) ? R(1) : R(0), X = pl;\n"
          << "        R const vpl = vp * pl;\n";
        int level = 0;
        int m = r.m0;
        for (int l = 0; l <= r.lmax; ++l)
            for (; m < r.lvEnd[l]; ++m)
            {
                MemberDesc const& md = F.member[g.mem0 + m];
                for (; level < md.level; ++level)
                    o << "        P *= X; X *= X;\n";
                o << "        acc[" << m << "] += vpl * P;\n";
            }
        for (; m < r.mend; ++m)
        {
            MemberDesc const& md = F.member[g.mem0 + m];
            o << "        { R const pw_ = ";
            if (md.useint) o << "pow_int(pl, " << md.zint - 1 << ")";
            else o << "pow(pl, " << lit(md.zeta - 1.0) << ")";
            o << "; acc[" << m << "] += pl > R(0) ? vpl * pw_ : R(0); }\n";
        }
        o << "        }\n";
    }
    o << "        break;\n    }\n";
}

int plan_maxm(ModelTables const& T)
{
    return T.maxGroupMembers <= 4 ? 4 : T.maxGroupMembers <= 8 ? 8 : 16;
}

int plan_cut(ModelTables const& T)
{
    int const c = T.cutType;
    return (c == CT_COS || c == CT_TANHU || c == CT_POLY2) ? c : -1;
}

// The waves-per-SIMD bounds of a tuning, with the built-in defaults for
// -1: FP32 chunks of at most 8 members run the SF kernel at seven waves
// (72 VGPRs; 3% faster than the compiler's choice on the 196,608-atom water
// box) and the force kernel at four or more; other forms leave it to the
// compiler.
int sf_waves(PlanTuning const& t, int fp32, int maxm)
{
    if (t.sfWaves >= 0) return t.sfWaves;
    return fp32 && maxm <= 8 ? 7 : 0;
}

int force_waves(PlanTuning const& t, int fp32, int maxm)
{
    if (t.forceWaves >= 0) return t.forceWaves;
    return fp32 && maxm <= 8 ? 4 : 0;
}

std::string waves_attr(int w)
{
    if (w <= 0) return "";
    return "__attribute__((amdgpu_waves_per_eu(" + std::to_string(w)
         + ", 8)))\n";
}

}  // namespace

// Kernel names carry the arithmetic, the member chunk and the waves bound,
// so that traces of different launch forms (a tuning sweep) keep them
// apart: hdnnp_plan_{sf,force}_{f32,f64}_m<chunk>_w<waves>.
std::string plan_kernel_name(ModelTables const& T, int fp32,
                             PlanTuning const& tune, bool force)
{
    int const maxm = plan_maxm(T);
    int const w = force ? force_waves(tune, fp32, maxm)
                        : sf_waves(tune, fp32, maxm);
    return std::string(force ? "hdnnp_plan_force" : "hdnnp_plan_sf")
         + (fp32 ? "_f32" : "_f64") + "_m" + std::to_string(maxm) + "_w"
         + std::to_string(w);
}

namespace {

std::string full_source(std::string const& plan, ModelTables const& T,
                        int fp32, PlanTuning const& tune)
{
    std::ostringstream o;
    char const* R = fp32 ? "float" : "double";
    int const maxm = plan_maxm(T);
    int const ct = plan_cut(T);
    // Models without weighted groups leave the weighted code out.
    o << "#define HDNNP_PLAN 1\n"
      << "#define HDNNP_WEIGHTED " << (T.weighted ? 1 : 0) << "\n"
      << "#include \"device_common.h\"\n"
      << "#include \"kernel_args.h\"\n"
      << "namespace hdnnp {\nnamespace {\n" << plan << "}\n}\n"
      << "#include \"force_core.h\"\n"
      << "#include \"sf_core.h\"\n"

      << "extern \"C\" __global__ __launch_bounds__(64)\n"
      << waves_attr(force_waves(tune, fp32, maxm))
      << "void " << plan_kernel_name(T, fp32, tune, true)
      << "(hdnnp::ForceArgs a)\n"
      << "{ hdnnp::force_backward_body<" << ct << ", false, " << maxm << ", "
      << R << ">(a); }\n"
      << "extern \"C\" __global__ __launch_bounds__(64)\n"
      << waves_attr(sf_waves(tune, fp32, maxm))
      << "void " << plan_kernel_name(T, fp32, tune, false)
      << "(hdnnp::SFArgs a)\n"
      << "{ hdnnp::sf_forward_body<" << maxm << ", " << ct << ", false, "
      << R << ">(a); }\n";
    return o.str();
}

struct Module
{
    hipModule_t mod = nullptr;
    PlanKernels k;
};

std::mutex g_mutex;
std::map<std::string, std::shared_ptr<Module>> g_cache;
bool g_warned = false;

bool verbose()
{
    char const* v = std::getenv("N2P2_AMD_RTC_VERBOSE");
    return v && *v && *v != '0';
}

// N2P2_AMD_RTC_DUMP=<directory>: the generated source of every compiled
// specialisation is also written there (hdnnp_plan_f32.hip /
// hdnnp_plan_f64.hip), for inspecting the code the kernels are built from.
void dump_source(std::string const& src, int fp32)
{
    char const* d = std::getenv("N2P2_AMD_RTC_DUMP");
    if (!d || !*d) return;
    std::string const path = std::string(d) + "/hdnnp_plan_"
                           + (fp32 ? "f32" : "f64") + ".hip";
    if (FILE* f = std::fopen(path.c_str(), "w"))
    {
        std::fwrite(src.data(), 1, src.size(), f);
        std::fclose(f);
    }
}

std::shared_ptr<Module> compile(std::string const& src, int fp32, int device,
                                std::string const& sfName,
                                std::string const& forceName)
{
    dump_source(src, fp32);
    auto t0 = std::chrono::steady_clock::now();
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "hdnnp_plan.hip", kRtcNumHeaders,
                            kRtcHeaders, kRtcHeaderNames) != HIPRTC_SUCCESS)
        throw std::runtime_error("hiprtcCreateProgram failed");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        throw std::runtime_error("hipGetDeviceProperties failed");
    std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
    // Packed FP32 instructions stay disabled, as for the ahead-of-time build.
    char const* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang",
                          "-target-feature", "-Xclang", "-packed-fp32-ops"};
    hiprtcResult const rc = hiprtcCompileProgram(prog, 7, opts);
    if (rc != HIPRTC_SUCCESS)
    {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        throw std::runtime_error("hipRTC compilation failed:\n" + log);
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    std::vector<char> code(n);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    auto m = std::make_shared<Module>();
    if (hipModuleLoadData(&m->mod, code.data()) != hipSuccess)
        throw std::runtime_error("hipModuleLoadData failed");
    if (hipModuleGetFunction(&m->k.force, m->mod, forceName.c_str())
            != hipSuccess
        || hipModuleGetFunction(&m->k.sf, m->mod, sfName.c_str())
            != hipSuccess)
        throw std::runtime_error("hipModuleGetFunction failed");

    if (verbose())
    {
        double const s = std::chrono::duration<double>(
            std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr, "[n2p2_amd] specialised kernels compiled in "
                             "%.2f s (%zu bytes)\n", s, n);
    }
    return m;
}

}  // namespace

std::string plan_source(Model::Flat const& F)
{
    std::ostringstream b, f;
    b << "template <typename R>\n"
         "__device__ __forceinline__ void plan_backward_members(\n"
         "    int gi, bool wide, R costijk, R r2sum, R rij, R rik, R rjk,\n"
         "    R invj, R invk, R invjk, R const* w, R const* wz, R& T1,\n"
         "    R& V, R& Ej, R& Ek, R& E3)\n{\n"
         "    (void)wide; (void)rij; (void)rik; (void)rjk; (void)invj;\n"
         "    (void)invk; (void)invjk; (void)r2sum;\n"
         "    switch (gi)\n    {\n";
    f << "template <typename R, int MAXM>\n"
         "__device__ __forceinline__ void plan_forward_members(\n"
         "    int gi, bool wide, R costijk, R pfc, R r2sum, R rij, R rik,\n"
         "    R rjk, R (&acc)[MAXM])\n{\n"
         "    (void)wide; (void)rij; (void)rik; (void)rjk; (void)r2sum;\n"
         "    switch (gi)\n    {\n";
    for (size_t gi = 0; gi < F.group.size(); ++gi)
    {
        GroupDesc const& g = F.group[gi];
        if (sf_radial(g.type)) continue;
        emit_backward_case(b, (int)gi, g, F);
        emit_forward_case(f, (int)gi, g, F);
    }
    b << "    default: break;\n    }\n}\n\n";
    f << "    default: break;\n    }\n}\n\n";
    return b.str() + f.str();
}

bool pair_record_enabled_by_default()
{
    char const* v = std::getenv("N2P2_AMD_PAIR_RECORD");
    return !(v && std::strcmp(v, "0") == 0);
}

bool plan_enabled_by_default()
{
    char const* v = std::getenv("N2P2_AMD_RTC");
    return !(v && std::strcmp(v, "0") == 0);
}

PlanKernels plan_kernels(Model::Flat const& F, ModelTables const& T,
                         int fp32, int device, PlanTuning const& tune)
{
    std::string const src = full_source(plan_source(F), T, fp32, tune);
    // Keyed by the full source (a few tens of KB): no collisions.
    std::string const key = std::to_string(device) + ":" + src;
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) return it->second ? it->second->k : PlanKernels();
    std::shared_ptr<Module> m;
    try
    {
        m = compile(src, fp32, device, plan_kernel_name(T, fp32, tune, false),
                    plan_kernel_name(T, fp32, tune, true));
    }
    catch (std::exception const& e)
    {
        if (!g_warned)
            std::fprintf(stderr, "[n2p2_amd] model-specialised kernels "
                                 "unavailable, using the generic kernels: "
                                 "%s\n", e.what());
        g_warned = true;
    }
    g_cache[key] = m;
    return m ? m->k : PlanKernels();
}

}  // namespace hdnnp
