// ml_internal.h — library-internal state behind the opaque C-ABI handles:
// the device context, the nonlinear (H, Theta) operator with its MG
// hierarchy, and the fused FAS engine used by the whole-solve entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

#include <memory>
#include <string>
#include <vector>

#include "ml_common.h"
#include "ml_kernels.h"
#include "mlhip.h"

namespace ml {

// Job tables of the batched calls, sent through pinned staging memory: two
// pinned buffers used in turn (the host waits only for the copy out of the
// buffer it refills) and one device buffer, reused in stream order (a copy
// into it runs after every kernel enqueued before it).
struct TableStage {
  void* dev = nullptr;
  size_t dev_cap = 0;
  void* host[2] = {nullptr, nullptr};
  size_t host_cap[2] = {0, 0};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int next = 0;
  hipError_t put(const void* src, size_t bytes, hipStream_t s, void** out);
  void release();
};

}  // namespace ml

struct ml_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  unsigned long long* d_norm = nullptr;   // device scratch for max-norms
  unsigned long long* h_norm = nullptr;   // pinned host mirror
  hipEvent_t events[16] = {};
  ml::TableStage tables;                    // job tables of the batched calls
  // launch timing of one kernel family (ml_ctx_timing_start/stop)
  int timing_kernel = 0;
  int timing_n = 0;
  long long timing_missed = 0;
  double timing_bytes = 0;
  int timing_cap = 0;                       // launches timed (<= the pool's size)
  std::vector<hipEvent_t> timing_start, timing_stop;
  ml::LaunchTimer timer(int kernel) {
    if (timing_kernel != kernel || timing_cap == 0) return ml::LaunchTimer{};
    return ml::LaunchTimer{timing_start.data(), timing_stop.data(), timing_cap, &timing_n,
                           &timing_bytes, &timing_missed};
  }
};

namespace ml {

// A kernel descriptor sent to the device through pinned staging memory.  The
// copy is skipped when the descriptor equals the last one sent (the device
// copy still holds it); otherwise it waits only for the previous copy out of
// the staging buffer (an event), not for the whole stream.
struct StagedDesc {
  std::vector<unsigned char> last;
  hipEvent_t ev = nullptr;
  hipError_t put(void* dev, void* pinned, const void* d, size_t bytes, hipStream_t s) {
    if (last.size() == bytes && std::memcmp(last.data(), d, bytes) == 0) return hipSuccess;
    last.clear();
    hipError_t e = ev ? hipEventSynchronize(ev) : hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return e;
    std::memcpy(pinned, d, bytes);
    if ((e = hipMemcpyAsync(dev, pinned, bytes, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
    if ((e = hipEventRecord(ev, s)) != hipSuccess) return e;
    const unsigned char* b = static_cast<const unsigned char*>(d);
    last.assign(b, b + bytes);
    return hipSuccess;
  }
  void release() {
    if (ev) hipEventDestroy(ev);
    ev = nullptr;
    last.clear();
  }
};

int fail(ml_ctx ctx, int code, const std::string& msg);
// device address of the no-allocation scratch array on `device` (the current
// device), looked up once per device (ml_noalloc.cpp)
int scratch_init(ml_ctx ctx, int device);
// largest dynamic LDS of one workgroup on the context's device (bytes)
int max_lds(ml_ctx ctx, size_t* out);
bool valid_grid(int dim, const int n[], const int periodic[], double dx);
bool odd_periodic_small_level(int dim, const int n[], const int periodic[]);

#define ML_HIP(ctx, expr)                                                                     \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      return ::ml::fail((ctx), ML_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
  } while (0)

struct HCLevel {
  Geom g;
  double* b[3] = {nullptr, nullptr, nullptr};
  double* a = nullptr;
  double* tmp = nullptr;   // ping-pong partner for the per-level relax entry point (lazy)
};

struct HostDriver;   // reference-API FAS driver state (ml_capi.cpp)
struct VCHostDriver; // reference-API linear MG driver state (ml_vc.cpp)
class HCEngine;      // fused FAS engine (ml_hc_engine.cpp)
struct AdjWork;      // adjoint-step workspace (ml_adjoint.hip)

}  // namespace ml

struct ml_hc_op_s {
  ml_ctx ctx = nullptr;
  int dim = 2;
  ml::ClosureP<double> cp{};
  ml::BCDesc bc{};
  ml_params params{};
  double alpha = 0.0, beta = -1.0;   // factory defaults (src/AMRNonLinearMultiCompOp.cpp:1622-1633)
  std::vector<ml::HCLevel> lev;
  std::unique_ptr<ml::HostDriver, void (*)(ml::HostDriver*)> host_driver{nullptr, nullptr};
  std::unique_ptr<ml::HCEngine, void (*)(ml::HCEngine*)> engine{nullptr, nullptr};
  std::unique_ptr<ml::AdjWork, void (*)(ml::AdjWork*)> adj{nullptr, nullptr};
  double* adv = nullptr;   // workspace of the advective source (allocated on first use)
  size_t adv_bytes = 0;

  ml::HCArgs<double> args(int depth) const {
    ml::HCArgs<double> A;
    A.g = lev[depth].g;
    A.cp = cp;
    A.bc = bc;
    A.alpha = alpha;
    A.beta = beta;
    A.a = lev[depth].a;
    for (int d = 0; d < 3; ++d) A.b[d] = lev[depth].b[d];
    return A;
  }
};

namespace ml {

// FAS V-cycles of the (H, Theta) operator with fused kernels: tiled levels
// above `lc`, one single-workgroup kernel for the levels from `lc` down.
// Same operation sequence (and FP64 results) as the reference-API driver.
class HCEngine {
 public:
  explicit HCEngine(ml_hc_op op);
  ~HCEngine();
  HCEngine(const HCEngine&) = delete;
  HCEngine& operator=(const HCEngine&) = delete;

  int solve(double* phi, const double* rhs, const ml_solve_opts& o, ml_solve_stats* st);
  int solve_be(double* phi_new, const double* phi_old, const double* src, double dt, const ml_solve_opts& o,
               ml_solve_stats* st);

 private:
  int init();
  int vcycle(double* phi, const double* rhs, const ml_solve_opts& o, const int* skip);
  int residual_norm(double* phi, const double* rhs, int slot, const int* skip, bool check);
  int read_norm(int slot, double* out);
  int read_ctl(SolveCtl* out);
  int write_ctl(const SolveCtl& c);
  int upload_desc(double* phi0, const double* rhs0, const ml_solve_opts& o);
  // FP32-refined form: FP64 defect correction around FP32 linear V-cycles
  int solve_refined(double* phi, const double* rhs, const ml_solve_opts& o, ml_solve_stats* st);
  int init_refined();
  int lin_vcycle(const ml_solve_opts& o);
  // state buffer pair of level l: {level-0 caller array or m_phi[l], m_alt[l]}
  void bufs(int l, double* phi, double* b[2]) const {
    b[0] = l == 0 ? phi : m_phi[l];
    b[1] = m_alt[l];
  }
  double* cur(int l, double* phi) const { return m_cur[l] ? m_alt[l] : (l == 0 ? phi : m_phi[l]); }

  // FP32-refined state per level; the corrections of tiled levels
  // ping-pong between m_lin[l].e and m_ealt[l] (m_ecur[l] says which)
  std::vector<LinLevel<float>> m_lin;
  std::vector<float*> m_ealt;
  std::vector<int> m_ecur;
  std::vector<float*> m_lin_owned;
  LinDesc* m_lindesc = nullptr;      // device
  LinDesc* m_hlindesc = nullptr;     // pinned host staging
  StagedDesc m_stage_lin;
  bool m_refined_ok = false;
  size_t m_lin_lds = 0;               // dynamic LDS of the coarse linear V-cycle

  ml_hc_op m_op;
  int m_L = 0, m_lc = 0;
  std::vector<double*> m_phi, m_rhs, m_phi0;   // per level (level 0: caller's arrays)
  std::vector<double*> m_alt;                  // relax ping-pong partners (tiled levels, and level 0)
  std::vector<int> m_cur;                      // which of the pair holds the state
  double* m_rhs_be = nullptr;                   // backward-Euler rhs of level 0
  unsigned long long* m_slots = nullptr;       // max-norm accumulators
  unsigned long long* m_hslots = nullptr;      // pinned mirror
  int m_nslots = 0;
  CoarseDesc<double>* m_desc = nullptr;        // device
  CoarseDesc<double>* m_hdesc = nullptr;       // pinned host staging
  StagedDesc m_stage_fas;
  SolveCtl* m_ctl = nullptr;                   // device iteration control
  unsigned int* m_arrive = nullptr;            // workgroup arrival counter of the fused stopping rule
  SolveCtl* m_hctl = nullptr;                  // pinned mirror
  int m_batch_fas = 4;                         // iterations enqueued per host check
  int m_batch_refined = 4;
  bool m_ok = false;
};

}  // namespace ml

// ---- linear variable-coefficient operator (DarcyBrinkmanOp / AMRProjectionOp)
namespace ml {
struct VCLevelStore {
  Geom g;
  double* a = nullptr;
  double* c = nullptr;
  double* b[3] = {nullptr, nullptr, nullptr};
  double* tmp = nullptr;               // partner for the per-level relax entry point (lazy)
};

// Workspace of one solve with an ml_vc_op's operator: per level the
// correction (with its ping-pong partner on tiled levels) and the residual
// right-hand side, the level-0 state partner, the backward-Euler rhs, norm
// slots, the coarse-kernel descriptor and the iteration control; and the
// stream the solve runs on (the context's for workspace 0, an own stream of
// default priority for the others, ordered after the context stream).
struct VCWork {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t joined = nullptr;          // recorded on `stream` when a solve ends
  std::vector<double*> e0, e1, r;       // per level
  double* r0alt = nullptr;              // level-0 residual partner (the fused residual writes it)
  double* alt0 = nullptr;
  double* rhs_be = nullptr;
  unsigned long long* slots = nullptr;
  VCDesc<double>* desc = nullptr;       // device
  VCDesc<double>* hdesc = nullptr;      // pinned staging
  StagedDesc stage;
  SolveCtl* ctl = nullptr;              // device iteration control
  SolveCtl* hctl = nullptr;             // pinned mirror
  unsigned int* arrive = nullptr;       // workgroup arrival counter of the fused stopping rule
  int batch = 4;                        // iterations enqueued per host check
  // two iterations (V-cycle + residual, even then odd) captured as a HIP
  // graph, and what the capture depends on
  hipGraphExec_t gexec = nullptr;
  const double* g_phi = nullptr;
  const double* g_rhs = nullptr;
  int g_nu1 = -1, g_nu2 = -1, g_bottom = -1, g_bsolver = -1, g_bimax = -1;
  double g_alpha = 0, g_beta = 0;
};
}  // namespace ml

struct ml_vc_op_s {
  ml_ctx ctx = nullptr;
  int dim = 2;
  ml::BCDesc bc{};
  double alpha = 1.0, beta = 1.0;   // factory defaults (src/DarcyBrinkmanOp.cpp:876-882)
  bool has_a = false, has_c = false;
  std::vector<ml::VCLevelStore> lev;
  int lc = 0;                       // first level handled by the single-workgroup kernel
  int nslots = 1024;                // norm slots per workspace (max_iter + 1)
  std::vector<ml::VCWork> work;     // solve workspaces; [0] runs on the context stream
  size_t coarse_lds = 0;            // dynamic LDS of the single-workgroup V-cycle
  size_t coarse_lds_bicg = 0;       // ... with the BiCGStab bottom (0: does not fit)
  // the reference-API operator objects and linear MG driver behind
  // ML_SOLVE_HOST_DRIVER (built at its first use)
  std::unique_ptr<ml::VCHostDriver, void (*)(ml::VCHostDriver*)> host_driver{nullptr, nullptr};

  ml::VCArgs<double> args(int depth) const {
    ml::VCArgs<double> A;
    A.g = lev[depth].g;
    A.bc = bc;
    A.alpha = alpha;
    A.beta = beta;
    A.a = has_a ? lev[depth].a : nullptr;
    A.c = has_c ? lev[depth].c : nullptr;
    for (int d = 0; d < 3; ++d) A.b[d] = lev[depth].b[d];
    return A;
  }
};

