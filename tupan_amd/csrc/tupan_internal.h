// Internal declarations shared by the translation units of libtupan_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/tupan_hip.h"

namespace tupan {

enum KernelId {
    K_PHI = TUPAN_K_PHI,
    K_ACC = TUPAN_K_ACC,
    K_ACC_JERK = TUPAN_K_ACC_JERK,
    K_SNAP_CRACKLE = TUPAN_K_SNAP_CRACKLE,
    K_TSTEP = TUPAN_K_TSTEP,
    K_NREG_X = TUPAN_K_NREG_X,
    K_NREG_V = TUPAN_K_NREG_V,
    K_PNACC = TUPAN_K_PNACC,
};
constexpr int kNumKernels = 8;

// Number of i/j input fields and of outputs of each pairwise kernel, in the
// reference's argument order (tupan/lib/extensions.py:100-387, 511-615).
inline int kernel_in_fields(int k)
{
    switch (k) {
    case K_PHI: return 5;
    case K_ACC: return 5;
    case K_ACC_JERK: return 8;
    case K_SNAP_CRACKLE: return 14;
    case K_TSTEP: return 8;
    case K_NREG_X: return 8;
    case K_NREG_V: return 7;
    case K_PNACC: return 8;
    default: return -1;
    }
}

inline int kernel_out_fields(int k)
{
    switch (k) {
    case K_PHI: return 1;
    case K_ACC: return 3;
    case K_ACC_JERK: return 6;
    case K_SNAP_CRACKLE: return 6;
    case K_TSTEP: return 2;
    case K_NREG_X: return 7;
    case K_NREG_V: return 4;
    case K_PNACC: return 3;
    default: return -1;
    }
}

// Per-device state.
struct Device {
    int id = -1;
    int cu_count = 0;
    int target_wg = 0;  // workgroups the launch planner aims for
    // scratch workspaces (j-split partials, near-pair lists), by slot
    void* ws[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t ws_bytes[4] = {0, 0, 0, 0};
    // the stream of each slot's last use, and an event that orders a use on
    // another stream after it
    hipStream_t ws_stream[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ws_used[4] = {false, false, false, false};
    hipEvent_t ws_event[4] = {nullptr, nullptr, nullptr, nullptr};
};

// State of the current device (initialised on first use).
const Device& current_device();

// Scratch workspace `slot` (0..3) of at least `bytes` on the current device,
// for work enqueued on `stream`.  Slots are independent buffers, so one call
// can hold several at once.  When the slot's last use was on another stream
// and neither stream is capturing a graph, `stream` is made to wait (no host
// synchronisation) for the work enqueued on that stream so far.  Growing a
// slot allocates a new buffer of at least twice the size and keeps the old
// one allocated until tupan_release_workspaces or the end of the process
// (captured HIP graphs may still address it).  Returns null on allocation
// failure (the old buffer then stays current).
void* workspace(size_t bytes, hipStream_t stream, int slot = 0);

// Error state (thread-local message).  set_error returns `code`.
int set_error(int code, const char* msg);
int check_hip(hipError_t e, const char* what);
int check_launch(const char* what);

// Optional HIP-event timing of the main (all-pairs) kernel of each call.
// timing_begin/timing_end bracket that launch on its stream when enabled.
void timing_begin(hipStream_t stream);
void timing_end(hipStream_t stream);

// Post-Newtonian parameters of PNAcc: the order and 1/c^k, k = 1..7 (the
// reference's CLIGHT, common.h:16-30).
struct PnParams {
    unsigned long order;
    double inv[7];
};

// One all-pairs call.  `pn` is required by K_PNACC and ignored otherwise.
template <class T>
int pairwise_call(int kernel, int mode, size_t ni, const T* const* ip, size_t nj,
                  const T* const* jp, T* const* out, T eta, void* stream,
                  const PnParams* pn = nullptr);

// pairwise_call (no PN parameters) with the j-split partial sums in the
// library's fixed device array: allocates nothing (pairwise.hip).
int pairwise_noalloc_call(int kernel, int mode, size_t ni, const double* const* ip, size_t nj,
                          const double* const* jp, double* const* out, double eta, void* stream);

// Tuning table of the pairwise kernels (pairwise.hip).
int tune_set(int kernel, int slot, int bucket, int shape, int wg_per_cu);
int tune_get(int kernel, int slot, int bucket, int* shape, int* wg_per_cu);
int tune_shape(int kernel, int shape, int* block, int* wpt, int* tile, int* unroll);

int hermite_step_call(int order, int adaptive, int mode, size_t n, void* const* arrays,
                      double* scratch, double* control, double eta, void* stream);
int sakura_step_call(int adaptive, int mode, size_t n, void* const* arrays, double* scratch,
                     double* ctl, double eta, int capacity, unsigned long long* counters,
                     int* status, void* stream);

int sakura_call(size_t ni, const double* const* ip, size_t nj, const double* const* jp,
                double dt, long flag, double* const* out, unsigned long long* counters,
                int* status, void* stream);

// Per-item log of the near-pair solves of the two-pass Sakura form.
struct SolveLogDev {
    float* features = nullptr;  // [4][capacity]: tau, sigma, kappa, eps
    float* target = nullptr;    // [capacity]: s r / dt - 1, NaN if unconverged
    int* iterations = nullptr;  // [capacity]: Laguerre steps of the pair
    int capacity = 0;
};

struct SakuraOptions {
    int path = 0;                 // 0 automatic, 1 one-pass, 2 two-pass
    const void* net = nullptr;    // packed Kepler-guess network (mlp.hip)
    SolveLogDev log;
    unsigned long long* fallbacks = nullptr;  // device counter of guess fallbacks
    int near_pairs = -1;          // out: near pairs listed (-1: one-pass or graph-safe form)
    int graph_capacity = -1;      // >= 0: two-pass form with no host synchronisation and a
                                  // near-pair list of this capacity (status bit 2 if exceeded)
    const double* dt_dev = nullptr;  // device dt (replaces the dt argument)
};

int sakura_call_ex(size_t ni, const double* const* ip, size_t nj, const double* const* jp,
                   double dt, long flag, double* const* out, SakuraOptions& opt,
                   unsigned long long* counters, int* status, void* stream);

// Device-driven shared-step SIA step ('s' and 'a' variants, sia.hip).
int sia_step_call(int adaptive, int acc_mode, int tstep_mode, size_t n, void* const* arrays,
                  double* scratch, double* control, double eta, int nops, const int* ops,
                  const double* coef_bridge, const double* coef_evolve, void* stream);

// Device-scheduled hierarchical SIA steps (sia.hip).
int sia_levels_create(int n, void* const* arrays, int narrays, const int* roles, int acc_mode,
                      int tstep_mode, double eta, int nbridge, const int* bridge_op,
                      const double* bridge_c, int nevolve, const int* evolve_op,
                      const double* evolve_c, void** plan_out);
int sia_levels_destroy(void* plan);
int sia_levels_set(void* plan, double t_curr, double t_end, double eta, int clear_pairs);
int sia_levels_enqueue(void* plan, int begin, int nslots, void* stream);
int sia_levels_state(void* plan, double* step, int* state, unsigned long long* pairs);

// Device SIA recursion for sets of at most 64 particles (sia.hip).
int sia_subtree_call(int n, void* const* arrays, int narrays, const int* roles, int acc_mode,
                     int tstep_mode, double eta, double tau, int nbridge, const int* bridge_op,
                     const double* bridge_c, int nevolve, const int* evolve_op,
                     const double* evolve_c, double* t_curr, unsigned long long* pairs,
                     int* status, void* stream);

// Kepler-guess network (mlp.hip).  mlp_infer: delta[k] for the features
// feat[4][n] (SoA) with the packed BF16 network `net`.
int mlp_infer(int n, const float* feat, const void* net, float* delta, hipStream_t stream);

int ensemble_plan_create(int nsys, const int* h_offsets, int min_lanes, void** plan_out);
int ensemble_plan_destroy(void* plan);
int ensemble_plan_info(const void* plan, int* nsys, int* ntot, int* nwaves);
int ensemble_plan_set_guess(void* plan, const void* net, unsigned long long* fallbacks);
int ensemble_sakura_call(const void* plan, const double* const* in8, double dt, long flag,
                         double* const* out6, unsigned long long* counters, int* status,
                         void* stream);
int ensemble_phi_call(const void* plan, const double* const* in8, double* phi, void* stream);

int kepler_pairs_call(size_t npairs, const int* idx, const double* m, double* rx, double* ry,
                      double* rz, const double* e2, double* vx, double* vy, double* vz,
                      double dt, int* status, void* stream);

int rsq_bias_probe(double* out);

}  // namespace tupan
