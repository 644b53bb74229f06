// C ABI of libtupan_hip.so: device state, scratch workspace, error state,
// the device-resident entry points and the reference-compatible host-pointer
// layer (tupan/lib/src/libtupan.h:2-246).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "tupan_internal.h"

namespace tupan {

namespace {

thread_local std::string g_error;
std::mutex g_dev_mutex;
std::vector<Device> g_devices;
std::vector<void*> g_stage;       // host-layer staging buffers, per device
std::vector<size_t> g_stage_bytes;
// outgrown workspaces (device, pointer), kept until released
std::vector<std::pair<int, void*>> g_retired;

Device& device_slot()
{
    int id = 0;
    if (hipGetDevice(&id) != hipSuccess) id = 0;
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    if (static_cast<int>(g_devices.size()) <= id) {
        g_devices.resize(id + 1);
        g_stage.resize(id + 1, nullptr);
        g_stage_bytes.resize(id + 1, 0);
    }
    Device& d = g_devices[id];
    if (d.id < 0) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, id) == hipSuccess) {
            d.cu_count = p.multiProcessorCount;
        } else {
            d.cu_count = 256;
        }
        // four 256-thread workgroups per CU keeps 16 waves per CU in flight
        d.target_wg = 4 * d.cu_count;
        d.id = id;
    }
    return d;
}

}  // namespace

const Device& current_device() { return device_slot(); }

namespace {

// Orders this use of a workspace slot after its last use when that was on
// another stream (event recorded there, waited on here).  Skipped while
// either stream is capturing: a capture must not overlap other streams' use
// of the library.
void order_slot(Device& d, int slot, hipStream_t stream)
{
    if (d.ws_used[slot] && d.ws_stream[slot] != stream) {
        hipStreamCaptureStatus here = hipStreamCaptureStatusNone;
        hipStreamCaptureStatus there = hipStreamCaptureStatusNone;
        bool ok = hipStreamIsCapturing(stream, &here) == hipSuccess &&
                  hipStreamIsCapturing(d.ws_stream[slot], &there) == hipSuccess;
        if (ok && here == hipStreamCaptureStatusNone && there == hipStreamCaptureStatusNone) {
            if (!d.ws_event[slot] &&
                hipEventCreateWithFlags(&d.ws_event[slot], hipEventDisableTiming) != hipSuccess) {
                d.ws_event[slot] = nullptr;
                ok = false;
            }
            ok = ok && hipEventRecord(d.ws_event[slot], d.ws_stream[slot]) == hipSuccess &&
                 hipStreamWaitEvent(stream, d.ws_event[slot], 0) == hipSuccess;
        }
        if (!ok) (void)hipGetLastError();
    }
    d.ws_stream[slot] = stream;
    d.ws_used[slot] = true;
}

}  // namespace

void* workspace(size_t bytes, hipStream_t stream, int slot)
{
    if (slot < 0 || slot > 3) return nullptr;
    Device& d = device_slot();
    order_slot(d, slot, stream);
    if (d.ws_bytes[slot] >= bytes) return d.ws[slot];
    // at least twice the old size, so a slot's retired buffers never add up
    // to more than its current one
    size_t want = bytes + bytes / 4;
    if (want < 2 * d.ws_bytes[slot]) want = 2 * d.ws_bytes[slot];
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (d.ws[slot]) {
        // Retired, not freed: a HIP graph captured earlier holds the old
        // address in its kernel nodes and may be replayed after this call
        // (and launches still in flight may use it), so it stays valid until
        // the process ends.
        std::lock_guard<std::mutex> lock(g_dev_mutex);
        g_retired.emplace_back(d.id, d.ws[slot]);
    }
    d.ws[slot] = p;
    d.ws_bytes[slot] = want;
    return p;
}

static void* staging(size_t bytes)
{
    Device& d = device_slot();
    const int id = d.id;
    if (g_stage_bytes[id] >= bytes) return g_stage[id];
    if (g_stage[id]) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_stage[id]);
        g_stage[id] = nullptr;
        g_stage_bytes[id] = 0;
    }
    size_t want = bytes + bytes / 4 + 4096;
    if (hipMalloc(&g_stage[id], want) != hipSuccess) {
        (void)hipGetLastError();
        g_stage[id] = nullptr;
        return nullptr;
    }
    g_stage_bytes[id] = want;
    return g_stage[id];
}

// ---------------------------------------------------------------------------
// Kernel timing: pairs of HIP events around each main kernel launch.
// ---------------------------------------------------------------------------
namespace {
bool g_timing = false;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_events;
size_t g_events_used = 0;
}  // namespace

void timing_begin(hipStream_t stream)
{
    if (!g_timing) return;
    if (g_events_used == g_events.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            g_timing = false;
            return;
        }
        g_events.emplace_back(a, b);
    }
    (void)hipEventRecord(g_events[g_events_used].first, stream);
}

void timing_end(hipStream_t stream)
{
    if (!g_timing || g_events_used >= g_events.size()) return;
    (void)hipEventRecord(g_events[g_events_used].second, stream);
    ++g_events_used;
}

int set_error(int code, const char* msg)
{
    g_error = msg;
    return code;
}

int check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return TUPAN_OK;
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return TUPAN_ERR_HIP;
}

int check_launch(const char* what) { return check_hip(hipGetLastError(), what); }

// ---------------------------------------------------------------------------
// Host-pointer layer: stage to the device, run, copy back, synchronously.
// ---------------------------------------------------------------------------
template <class T>
static void fail_outputs(const char* fn, T* const* out, const size_t* lens, int nout)
{
    fprintf(stderr, "tupan_hip: %s failed: %s\n", fn, g_error.c_str());
    for (int k = 0; k < nout; ++k)
        for (size_t i = 0; i < lens[k]; ++i) out[k][i] = static_cast<T>(NAN);
}

template <class T>
static void host_pairwise(const char* fn, int kernel, size_t ni, const T* const* ih, size_t nj,
                          const T* const* jh, T* const* oh, T eta,
                          const PnParams* pn = nullptr)
{
    const int nin = kernel_in_fields(kernel);
    const int nout = kernel_out_fields(kernel);
    size_t lens[8];
    for (int k = 0; k < nout; ++k) lens[k] = ni;
    if (ni == 0) return;
    const size_t bytes = sizeof(T) * (nin * (ni + nj) + nout * ni);
    T* base = static_cast<T*>(staging(bytes));
    if (!base) {
        set_error(TUPAN_ERR_ALLOC, "staging allocation failed");
        fail_outputs(fn, oh, lens, nout);
        return;
    }
    const T* ip[14];
    const T* jp[14];
    T* op[8];
    int rc = TUPAN_OK;
    T* cur = base;
    for (int f = 0; f < nin && rc == TUPAN_OK; ++f) {
        rc = check_hip(hipMemcpy(cur, ih[f], sizeof(T) * ni, hipMemcpyHostToDevice), "copy in");
        ip[f] = cur;
        cur += ni;
    }
    for (int f = 0; f < nin && rc == TUPAN_OK; ++f) {
        if (nj) rc = check_hip(hipMemcpy(cur, jh[f], sizeof(T) * nj, hipMemcpyHostToDevice), "copy in");
        jp[f] = cur;
        cur += nj;
    }
    for (int k = 0; k < nout; ++k) {
        op[k] = cur;
        cur += ni;
    }
    if (rc == TUPAN_OK)
        rc = pairwise_call<T>(kernel, TUPAN_MODE_FP64, ni, ip, nj, jp, op, eta, nullptr, pn);
    for (int k = 0; k < nout && rc == TUPAN_OK; ++k)
        rc = check_hip(hipMemcpy(oh[k], op[k], sizeof(T) * ni, hipMemcpyDeviceToHost), "copy out");
    if (rc != TUPAN_OK) fail_outputs(fn, oh, lens, nout);
}

static void host_sakura(size_t ni, const double* const* ih, size_t nj, const double* const* jh,
                        double dt, long flag, double* const* oh)
{
    size_t lens[6] = {ni, ni, ni, ni, ni, ni};
    if (ni == 0) return;
    const size_t bytes = sizeof(double) * (8 * (ni + nj) + 6 * ni) + 64;
    double* base = static_cast<double*>(staging(bytes));
    if (!base) {
        set_error(TUPAN_ERR_ALLOC, "staging allocation failed");
        fail_outputs("sakura_kernel", oh, lens, 6);
        return;
    }
    const double* ip[8];
    const double* jp[8];
    double* op[6];
    int rc = TUPAN_OK;
    double* cur = base;
    int* status = reinterpret_cast<int*>(cur);
    cur += 8;
    for (int f = 0; f < 8 && rc == TUPAN_OK; ++f) {
        rc = check_hip(hipMemcpy(cur, ih[f], 8 * ni, hipMemcpyHostToDevice), "copy in");
        ip[f] = cur;
        cur += ni;
    }
    for (int f = 0; f < 8 && rc == TUPAN_OK; ++f) {
        if (nj) rc = check_hip(hipMemcpy(cur, jh[f], 8 * nj, hipMemcpyHostToDevice), "copy in");
        jp[f] = cur;
        cur += nj;
    }
    for (int k = 0; k < 6; ++k) {
        op[k] = cur;
        cur += ni;
    }
    if (rc == TUPAN_OK) rc = check_hip(hipMemset(status, 0, sizeof(int)), "memset");
    if (rc == TUPAN_OK) rc = sakura_call(ni, ip, nj, jp, dt, flag, op, nullptr, status, nullptr);
    int st = 0;
    if (rc == TUPAN_OK)
        rc = check_hip(hipMemcpy(&st, status, sizeof(int), hipMemcpyDeviceToHost), "copy out");
    if (rc == TUPAN_OK && st != 0)
        rc = set_error(TUPAN_ERR_SOLVER, "Kepler solver exhausted its bounded step subdivision");
    for (int k = 0; k < 6 && rc == TUPAN_OK; ++k)
        rc = check_hip(hipMemcpy(oh[k], op[k], 8 * ni, hipMemcpyDeviceToHost), "copy out");
    if (rc != TUPAN_OK) fail_outputs("sakura_kernel", oh, lens, 6);
}

}  // namespace tupan

using namespace tupan;

extern "C" {

const char* tupan_last_error(void) { return g_error.c_str(); }

int tupan_version(void) { return 100; }

int tupan_release_workspaces(void)
{
    Device& d = device_slot();
    int rc = check_hip(hipDeviceSynchronize(), "synchronise before release");
    if (rc != TUPAN_OK) return rc;
    for (int k = 0; k < 4; ++k) {
        if (d.ws[k]) (void)hipFree(d.ws[k]);
        d.ws[k] = nullptr;
        d.ws_bytes[k] = 0;
        d.ws_used[k] = false;
        d.ws_stream[k] = nullptr;
    }
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    std::vector<std::pair<int, void*>> keep;
    for (const auto& r : g_retired) {
        if (r.first == d.id)
            (void)hipFree(r.second);
        else
            keep.push_back(r);
    }
    g_retired.swap(keep);
    if (g_stage[d.id]) {
        (void)hipFree(g_stage[d.id]);
        g_stage[d.id] = nullptr;
        g_stage_bytes[d.id] = 0;
    }
    return check_hip(hipGetLastError(), "release");
}

int tupan_timing(int enable)
{
    g_timing = enable != 0;
    g_events_used = 0;
    return TUPAN_OK;
}

int tupan_timing_read(double* total_ms, int* count)
{
    double tot = 0.0;
    for (size_t k = 0; k < g_events_used; ++k) {
        int rc = check_hip(hipEventSynchronize(g_events[k].second), "event sync");
        if (rc != TUPAN_OK) return rc;
        float ms = 0.0f;
        rc = check_hip(hipEventElapsedTime(&ms, g_events[k].first, g_events[k].second),
                       "event time");
        if (rc != TUPAN_OK) return rc;
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = static_cast<int>(g_events_used);
    return TUPAN_OK;
}

int tupan_timing_read_each(double* ms, int max, int* count)
{
    const int n = static_cast<int>(g_events_used);
    if (count) *count = n;
    if (!ms && max > 0) return set_error(TUPAN_ERR_ARG, "null output");
    for (int k = 0; k < n && k < max; ++k) {
        int rc = check_hip(hipEventSynchronize(g_events[k].second), "event sync");
        if (rc != TUPAN_OK) return rc;
        float t = 0.0f;
        rc = check_hip(hipEventElapsedTime(&t, g_events[k].first, g_events[k].second),
                       "event time");
        if (rc != TUPAN_OK) return rc;
        ms[k] = t;
    }
    return TUPAN_OK;
}

int tupan_tune_set(int kernel, int slot, int bucket, int shape, int wg_per_cu)
{
    return tune_set(kernel, slot, bucket, shape, wg_per_cu);
}

int tupan_tune_get(int kernel, int slot, int bucket, int* shape, int* wg_per_cu)
{
    return tune_get(kernel, slot, bucket, shape, wg_per_cu);
}

int tupan_tune_shape(int kernel, int shape, int* block, int* wpt, int* tile, int* unroll)
{
    return tune_shape(kernel, shape, block, wpt, tile, unroll);
}

int tupan_rsq_bias(double* out)
{
    if (!out) return set_error(TUPAN_ERR_ARG, "null output");
    return rsq_bias_probe(out);
}

int tupan_device_info(int* cu_count, int* target_workgroups)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        if (cu_count) *cu_count = -1;
        if (target_workgroups) *target_workgroups = -1;
        return set_error(TUPAN_ERR_HIP, "no HIP device");
    }
    const Device& d = current_device();
    if (cu_count) *cu_count = d.cu_count;
    if (target_workgroups) *target_workgroups = d.target_wg;
    return TUPAN_OK;
}

// ---- host-pointer layer (reference ABI, FP64) ------------------------------

void phi_kernel(const unsigned long ni, const double* _im, const double* _irx,
                const double* _iry, const double* _irz, const double* _ie2,
                const unsigned long nj, const double* _jm, const double* _jrx,
                const double* _jry, const double* _jrz, const double* _je2, double* _iphi)
{
    const double* ih[5] = {_im, _irx, _iry, _irz, _ie2};
    const double* jh[5] = {_jm, _jrx, _jry, _jrz, _je2};
    double* oh[1] = {_iphi};
    host_pairwise<double>("phi_kernel", K_PHI, ni, ih, nj, jh, oh, 0.0);
}

void acc_kernel(const unsigned long ni, const double* _im, const double* _irx,
                const double* _iry, const double* _irz, const double* _ie2,
                const unsigned long nj, const double* _jm, const double* _jrx,
                const double* _jry, const double* _jrz, const double* _je2, double* _iax,
                double* _iay, double* _iaz)
{
    const double* ih[5] = {_im, _irx, _iry, _irz, _ie2};
    const double* jh[5] = {_jm, _jrx, _jry, _jrz, _je2};
    double* oh[3] = {_iax, _iay, _iaz};
    host_pairwise<double>("acc_kernel", K_ACC, ni, ih, nj, jh, oh, 0.0);
}

void acc_jerk_kernel(const unsigned long ni, const double* _im, const double* _irx,
                     const double* _iry, const double* _irz, const double* _ie2,
                     const double* _ivx, const double* _ivy, const double* _ivz,
                     const unsigned long nj, const double* _jm, const double* _jrx,
                     const double* _jry, const double* _jrz, const double* _je2,
                     const double* _jvx, const double* _jvy, const double* _jvz, double* _iax,
                     double* _iay, double* _iaz, double* _ijx, double* _ijy, double* _ijz)
{
    const double* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const double* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    double* oh[6] = {_iax, _iay, _iaz, _ijx, _ijy, _ijz};
    host_pairwise<double>("acc_jerk_kernel", K_ACC_JERK, ni, ih, nj, jh, oh, 0.0);
}

void snap_crackle_kernel(const unsigned long ni, const double* _im, const double* _irx,
                         const double* _iry, const double* _irz, const double* _ie2,
                         const double* _ivx, const double* _ivy, const double* _ivz,
                         const double* _iax, const double* _iay, const double* _iaz,
                         const double* _ijx, const double* _ijy, const double* _ijz,
                         const unsigned long nj, const double* _jm, const double* _jrx,
                         const double* _jry, const double* _jrz, const double* _je2,
                         const double* _jvx, const double* _jvy, const double* _jvz,
                         const double* _jax, const double* _jay, const double* _jaz,
                         const double* _jjx, const double* _jjy, const double* _jjz,
                         double* _isx, double* _isy, double* _isz, double* _icx, double* _icy,
                         double* _icz)
{
    const double* ih[14] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy,
                            _ivz, _iax, _iay, _iaz, _ijx, _ijy, _ijz};
    const double* jh[14] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy,
                            _jvz, _jax, _jay, _jaz, _jjx, _jjy, _jjz};
    double* oh[6] = {_isx, _isy, _isz, _icx, _icy, _icz};
    host_pairwise<double>("snap_crackle_kernel", K_SNAP_CRACKLE, ni, ih, nj, jh, oh, 0.0);
}

void tstep_kernel(const unsigned long ni, const double* _im, const double* _irx,
                  const double* _iry, const double* _irz, const double* _ie2,
                  const double* _ivx, const double* _ivy, const double* _ivz,
                  const unsigned long nj, const double* _jm, const double* _jrx,
                  const double* _jry, const double* _jrz, const double* _je2,
                  const double* _jvx, const double* _jvy, const double* _jvz, const double eta,
                  double* _idt_a, double* _idt_b)
{
    const double* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const double* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    double* oh[2] = {_idt_a, _idt_b};
    host_pairwise<double>("tstep_kernel", K_TSTEP, ni, ih, nj, jh, oh, eta);
}

void nreg_Xkernel(const unsigned long ni, const double* _im, const double* _irx,
                  const double* _iry, const double* _irz, const double* _ie2,
                  const double* _ivx, const double* _ivy, const double* _ivz,
                  const unsigned long nj, const double* _jm, const double* _jrx,
                  const double* _jry, const double* _jrz, const double* _je2,
                  const double* _jvx, const double* _jvy, const double* _jvz, const double dt,
                  double* _idrx, double* _idry, double* _idrz, double* _iax, double* _iay,
                  double* _iaz, double* _iu)
{
    const double* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const double* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    double* oh[7] = {_idrx, _idry, _idrz, _iax, _iay, _iaz, _iu};
    host_pairwise<double>("nreg_Xkernel", K_NREG_X, ni, ih, nj, jh, oh, dt);
}

void nreg_Vkernel(const unsigned long ni, const double* _im, const double* _ivx,
                  const double* _ivy, const double* _ivz, const double* _iax,
                  const double* _iay, const double* _iaz, const unsigned long nj,
                  const double* _jm, const double* _jvx, const double* _jvy, const double* _jvz,
                  const double* _jax, const double* _jay, const double* _jaz, const double dt,
                  double* _idvx, double* _idvy, double* _idvz, double* _ik)
{
    const double* ih[7] = {_im, _ivx, _ivy, _ivz, _iax, _iay, _iaz};
    const double* jh[7] = {_jm, _jvx, _jvy, _jvz, _jax, _jay, _jaz};
    double* oh[4] = {_idvx, _idvy, _idvz, _ik};
    host_pairwise<double>("nreg_Vkernel", K_NREG_V, ni, ih, nj, jh, oh, dt);
}

void pnacc_kernel(const unsigned long ni, const double* _im, const double* _irx,
                  const double* _iry, const double* _irz, const double* _ie2,
                  const double* _ivx, const double* _ivy, const double* _ivz,
                  const unsigned long nj, const double* _jm, const double* _jrx,
                  const double* _jry, const double* _jrz, const double* _je2,
                  const double* _jvx, const double* _jvy, const double* _jvz,
                  const unsigned long order, const double inv1, const double inv2,
                  const double inv3, const double inv4, const double inv5, const double inv6,
                  const double inv7, double* _ipnax, double* _ipnay, double* _ipnaz)
{
    const double* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const double* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    double* oh[3] = {_ipnax, _ipnay, _ipnaz};
    const PnParams pn = {order, {inv1, inv2, inv3, inv4, inv5, inv6, inv7}};
    host_pairwise<double>("pnacc_kernel", K_PNACC, ni, ih, nj, jh, oh, 0.0, &pn);
}

void sakura_kernel(const unsigned long ni, const double* _im, const double* _irx,
                   const double* _iry, const double* _irz, const double* _ie2,
                   const double* _ivx, const double* _ivy, const double* _ivz,
                   const unsigned long nj, const double* _jm, const double* _jrx,
                   const double* _jry, const double* _jrz, const double* _je2,
                   const double* _jvx, const double* _jvy, const double* _jvz, const double dt,
                   const long flag, double* _idrx, double* _idry, double* _idrz,
                   double* _idvx, double* _idvy, double* _idvz)
{
    const double* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const double* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    double* oh[6] = {_idrx, _idry, _idrz, _idvx, _idvy, _idvz};
    host_sakura(ni, ih, nj, jh, dt, flag, oh);
}

void kepler_solver_kernel(const double* _im, const double* _irx, const double* _iry,
                          const double* _irz, const double* _ie2, const double* _ivx,
                          const double* _ivy, const double* _ivz, const double dt,
                          double* _ir1x, double* _ir1y, double* _ir1z, double* _iv1x,
                          double* _iv1y, double* _iv1z)
{
    // stage: m rx ry rz e2 vx vy vz (2 each), pair index, status
    double* outs[6] = {_ir1x, _ir1y, _ir1z, _iv1x, _iv1y, _iv1z};
    size_t lens[6] = {2, 2, 2, 2, 2, 2};
    double* base = static_cast<double*>(staging(sizeof(double) * 16 + 64));
    if (!base) {
        set_error(TUPAN_ERR_ALLOC, "staging allocation failed");
        fail_outputs("kepler_solver_kernel", outs, lens, 6);
        return;
    }
    const double* in[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    double host[16];
    for (int f = 0; f < 8; ++f) {
        host[2 * f] = in[f][0];
        host[2 * f + 1] = in[f][1];
    }
    int* meta = reinterpret_cast<int*>(base + 16);
    int hmeta[3] = {0, 1, 0};
    int rc = check_hip(hipMemcpy(base, host, sizeof(host), hipMemcpyHostToDevice), "copy in");
    if (rc == TUPAN_OK)
        rc = check_hip(hipMemcpy(meta, hmeta, sizeof(hmeta), hipMemcpyHostToDevice), "copy in");
    if (rc == TUPAN_OK)
        rc = kepler_pairs_call(1, meta, base + 0, base + 2, base + 4, base + 6, base + 8,
                               base + 10, base + 12, base + 14, dt, meta + 2, nullptr);
    if (rc == TUPAN_OK)
        rc = check_hip(hipMemcpy(host, base, sizeof(host), hipMemcpyDeviceToHost), "copy out");
    if (rc == TUPAN_OK)
        rc = check_hip(hipMemcpy(hmeta, meta, sizeof(hmeta), hipMemcpyDeviceToHost), "copy out");
    if (rc == TUPAN_OK && hmeta[2] != 0)
        rc = set_error(TUPAN_ERR_SOLVER, "Kepler solver exhausted its bounded step subdivision");
    if (rc != TUPAN_OK) {
        fail_outputs("kepler_solver_kernel", outs, lens, 6);
        return;
    }
    const int map[6] = {1, 2, 3, 5, 6, 7};  // rx ry rz vx vy vz
    for (int k = 0; k < 6; ++k) {
        outs[k][0] = host[2 * map[k]];
        outs[k][1] = host[2 * map[k] + 1];
    }
}

// ---- host layer of the float build (used by libtupan_hip_f32.so) ----------

int tupan_host_pairwise_f32(int kernel, size_t ni, const float* const* ih, size_t nj,
                            const float* const* jh, float* const* oh, float eta)
{
    static const char* names[kNumKernels] = {"phi_kernel",          "acc_kernel",
                                             "acc_jerk_kernel",     "snap_crackle_kernel",
                                             "tstep_kernel",        "nreg_Xkernel",
                                             "nreg_Vkernel",        "pnacc_kernel"};
    if (kernel < 0 || kernel >= kNumKernels || kernel == K_PNACC)
        return set_error(TUPAN_ERR_ARG, "unknown pairwise kernel");
    g_error.clear();
    host_pairwise<float>(names[kernel], kernel, ni, ih, nj, jh, oh, eta);
    return g_error.empty() ? TUPAN_OK : TUPAN_ERR_HIP;
}

// Sakura and Kepler of the float build run in FP64 (the device solver is
// FP64 only) and round the results to float.
int tupan_host_pnacc_f32(size_t ni, const float* const* ih, size_t nj, const float* const* jh,
                         unsigned long order, const float* inv7, float* const* oh)
{
    if (!inv7) return set_error(TUPAN_ERR_ARG, "null 1/c^k array");
    PnParams pn = {order, {}};
    for (int k = 0; k < 7; ++k) pn.inv[k] = inv7[k];
    g_error.clear();
    host_pairwise<float>("pnacc_kernel", K_PNACC, ni, ih, nj, jh, oh, 0.0f, &pn);
    return g_error.empty() ? TUPAN_OK : TUPAN_ERR_HIP;
}

int tupan_host_sakura_f32(size_t ni, const float* const* ih, size_t nj, const float* const* jh,
                          float dt, int flag, float* const* oh)
{
    std::vector<double> buf(8 * (ni + nj) + 6 * ni);
    const double* ip[8];
    const double* jp[8];
    double* op[6];
    double* cur = buf.data();
    for (int f = 0; f < 8; ++f) {
        for (size_t i = 0; i < ni; ++i) cur[i] = ih[f][i];
        ip[f] = cur;
        cur += ni;
    }
    for (int f = 0; f < 8; ++f) {
        for (size_t j = 0; j < nj; ++j) cur[j] = jh[f][j];
        jp[f] = cur;
        cur += nj;
    }
    for (int k = 0; k < 6; ++k) {
        op[k] = cur;
        cur += ni;
    }
    g_error.clear();
    host_sakura(ni, ip, nj, jp, static_cast<double>(dt), static_cast<long>(flag), op);
    for (int k = 0; k < 6; ++k)
        for (size_t i = 0; i < ni; ++i) oh[k][i] = static_cast<float>(op[k][i]);
    return g_error.empty() ? TUPAN_OK : TUPAN_ERR_HIP;
}

int tupan_host_kepler_f32(const float* const* in8, float dt, float* const* out6)
{
    double in[8][2], o[6][2];
    const double* ip[8];
    double* op[6];
    for (int f = 0; f < 8; ++f) {
        in[f][0] = in8[f][0];
        in[f][1] = in8[f][1];
        ip[f] = in[f];
    }
    for (int k = 0; k < 6; ++k) op[k] = o[k];
    g_error.clear();
    kepler_solver_kernel(ip[0], ip[1], ip[2], ip[3], ip[4], ip[5], ip[6], ip[7],
                         static_cast<double>(dt), op[0], op[1], op[2], op[3], op[4], op[5]);
    for (int k = 0; k < 6; ++k) {
        out6[k][0] = static_cast<float>(o[k][0]);
        out6[k][1] = static_cast<float>(o[k][1]);
    }
    return g_error.empty() ? TUPAN_OK : TUPAN_ERR_HIP;
}

// ---- device-resident layer --------------------------------------------------

int tupan_pairwise(int kernel, int mode, size_t ni, const double* const* ifields, size_t nj,
                   const double* const* jfields, double* const* out, double eta, void* stream)
{
    return pairwise_call<double>(kernel, mode, ni, ifields, nj, jfields, out, eta, stream);
}

int tupan_pairwise_noalloc(int kernel, int mode, size_t ni, const double* const* ifields,
                           size_t nj, const double* const* jfields, double* const* out,
                           double eta, void* stream)
{
    return pairwise_noalloc_call(kernel, mode, ni, ifields, nj, jfields, out, eta, stream);
}

int tupan_pairwise_f32(int kernel, int mode, size_t ni, const float* const* ifields, size_t nj,
                       const float* const* jfields, float* const* out, float eta, void* stream)
{
    return pairwise_call<float>(kernel, mode, ni, ifields, nj, jfields, out, eta, stream);
}

int tupan_hermite_step(int order, int adaptive, int mode, size_t n, void* const* arrays,
                       double* scratch, double* control, double eta, void* stream)
{
    return hermite_step_call(order, adaptive, mode, n, arrays, scratch, control, eta, stream);
}

int tupan_sia_step(int adaptive, int acc_mode, int tstep_mode, size_t n, void* const* arrays,
                   double* scratch, double* control, double eta, int nops, const int* ops,
                   const double* coef_bridge, const double* coef_evolve, void* stream)
{
    return sia_step_call(adaptive, acc_mode, tstep_mode, n, arrays, scratch, control, eta, nops,
                         ops, coef_bridge, coef_evolve, stream);
}

int tupan_sia_levels_create(int n, void* const* arrays, int narrays, const int* roles,
                            int acc_mode, int tstep_mode, double eta, int nbridge,
                            const int* bridge_op, const double* bridge_c, int nevolve,
                            const int* evolve_op, const double* evolve_c,
                            tupan_sia_levels** plan)
{
    void* p = nullptr;
    const int rc = sia_levels_create(n, arrays, narrays, roles, acc_mode, tstep_mode, eta, nbridge,
                                     bridge_op, bridge_c, nevolve, evolve_op, evolve_c, &p);
    if (plan) *plan = static_cast<tupan_sia_levels*>(p);
    return rc;
}

int tupan_sia_levels_destroy(tupan_sia_levels* plan) { return sia_levels_destroy(plan); }

int tupan_sia_levels_set(tupan_sia_levels* plan, double t_curr, double t_end, double eta,
                         int clear_pairs)
{
    return sia_levels_set(plan, t_curr, t_end, eta, clear_pairs);
}

int tupan_sia_levels_enqueue(tupan_sia_levels* plan, int begin, int nslots, void* stream)
{
    return sia_levels_enqueue(plan, begin, nslots, stream);
}

int tupan_sia_levels_state(tupan_sia_levels* plan, double* step, int* state,
                           unsigned long long* pairs)
{
    return sia_levels_state(plan, step, state, pairs);
}

int tupan_pnacc(int mode, size_t ni, const double* const* ifields, size_t nj,
                const double* const* jfields, unsigned long order, const double* inv7,
                double* const* out, void* stream)
{
    if (!inv7) return set_error(TUPAN_ERR_ARG, "null 1/c^k array");
    PnParams pn = {order, {}};
    for (int k = 0; k < 7; ++k) pn.inv[k] = inv7[k];
    return pairwise_call<double>(K_PNACC, mode, ni, ifields, nj, jfields, out, 0.0, stream, &pn);
}

int tupan_pnacc_f32(int mode, size_t ni, const float* const* ifields, size_t nj,
                    const float* const* jfields, unsigned long order, const float* inv7,
                    float* const* out, void* stream)
{
    if (!inv7) return set_error(TUPAN_ERR_ARG, "null 1/c^k array");
    PnParams pn = {order, {}};
    for (int k = 0; k < 7; ++k) pn.inv[k] = inv7[k];
    return pairwise_call<float>(K_PNACC, mode, ni, ifields, nj, jfields, out, 0.0f, stream, &pn);
}

int tupan_phi(int mode, size_t ni, const double* im, const double* irx, const double* iry,
              const double* irz, const double* ie2, size_t nj, const double* jm,
              const double* jrx, const double* jry, const double* jrz, const double* je2,
              double* iphi, void* stream)
{
    const double* ip[5] = {im, irx, iry, irz, ie2};
    const double* jp[5] = {jm, jrx, jry, jrz, je2};
    double* op[1] = {iphi};
    return pairwise_call<double>(K_PHI, mode, ni, ip, nj, jp, op, 0.0, stream);
}

int tupan_acc(int mode, size_t ni, const double* im, const double* irx, const double* iry,
              const double* irz, const double* ie2, size_t nj, const double* jm,
              const double* jrx, const double* jry, const double* jrz, const double* je2,
              double* iax, double* iay, double* iaz, void* stream)
{
    const double* ip[5] = {im, irx, iry, irz, ie2};
    const double* jp[5] = {jm, jrx, jry, jrz, je2};
    double* op[3] = {iax, iay, iaz};
    return pairwise_call<double>(K_ACC, mode, ni, ip, nj, jp, op, 0.0, stream);
}

int tupan_acc_jerk(int mode, size_t ni, const double* im, const double* irx, const double* iry,
                   const double* irz, const double* ie2, const double* ivx, const double* ivy,
                   const double* ivz, size_t nj, const double* jm, const double* jrx,
                   const double* jry, const double* jrz, const double* je2, const double* jvx,
                   const double* jvy, const double* jvz, double* iax, double* iay, double* iaz,
                   double* ijx, double* ijy, double* ijz, void* stream)
{
    const double* ip[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    const double* jp[8] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz};
    double* op[6] = {iax, iay, iaz, ijx, ijy, ijz};
    return pairwise_call<double>(K_ACC_JERK, mode, ni, ip, nj, jp, op, 0.0, stream);
}

int tupan_snap_crackle(int mode, size_t ni, const double* const* ifields, size_t nj,
                       const double* const* jfields, double* isx, double* isy, double* isz,
                       double* icx, double* icy, double* icz, void* stream)
{
    double* op[6] = {isx, isy, isz, icx, icy, icz};
    return pairwise_call<double>(K_SNAP_CRACKLE, mode, ni, ifields, nj, jfields, op, 0.0,
                                 stream);
}

int tupan_tstep(int mode, size_t ni, const double* im, const double* irx, const double* iry,
                const double* irz, const double* ie2, const double* ivx, const double* ivy,
                const double* ivz, size_t nj, const double* jm, const double* jrx,
                const double* jry, const double* jrz, const double* je2, const double* jvx,
                const double* jvy, const double* jvz, double eta, double* idt_a, double* idt_b,
                void* stream)
{
    const double* ip[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    const double* jp[8] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz};
    double* op[2] = {idt_a, idt_b};
    return pairwise_call<double>(K_TSTEP, mode, ni, ip, nj, jp, op, eta, stream);
}

int tupan_sakura(size_t ni, const double* im, const double* irx, const double* iry,
                 const double* irz, const double* ie2, const double* ivx, const double* ivy,
                 const double* ivz, size_t nj, const double* jm, const double* jrx,
                 const double* jry, const double* jrz, const double* je2, const double* jvx,
                 const double* jvy, const double* jvz, double dt, long flag, double* idrx,
                 double* idry, double* idrz, double* idvx, double* idvy, double* idvz,
                 unsigned long long* counters, int* status, void* stream)
{
    const double* ip[8] = {im, irx, iry, irz, ie2, ivx, ivy, ivz};
    const double* jp[8] = {jm, jrx, jry, jrz, je2, jvx, jvy, jvz};
    double* op[6] = {idrx, idry, idrz, idvx, idvy, idvz};
    return sakura_call(ni, ip, nj, jp, dt, flag, op, counters, status, stream);
}

int tupan_sakura_ex(size_t ni, const double* const* ifields, size_t nj,
                    const double* const* jfields, double dt, long flag, double* const* out,
                    int path, const void* net, const tupan_solve_log* log,
                    unsigned long long* counters, int* near_pairs, int* status, void* stream)
{
    if (path < 0 || path > 2) return set_error(TUPAN_ERR_ARG, "path must be 0, 1 or 2");
    if (!ifields || !jfields || !out) return set_error(TUPAN_ERR_ARG, "null field list");
    if (reinterpret_cast<uintptr_t>(net) % 16 != 0)
        return set_error(TUPAN_ERR_ARG, "the packed network must be 16-byte aligned");
    SakuraOptions opt;
    opt.path = path;
    opt.net = net;
    if (log) {
        if (log->capacity < 0 || (log->capacity > 0 &&
                                  (!log->features || !log->target || !log->iterations)))
            return set_error(TUPAN_ERR_ARG, "bad solve log");
        if (log->capacity > 0) {
            opt.log.features = log->features;
            opt.log.target = log->target;
            opt.log.iterations = log->iterations;
            opt.log.capacity = log->capacity;
        }
    }
    opt.fallbacks = counters ? counters + 4 : nullptr;
    const int rc = sakura_call_ex(ni, ifields, nj, jfields, dt, flag, out, opt, counters,
                                  status, stream);
    if (near_pairs) *near_pairs = opt.near_pairs;
    return rc;
}

int tupan_sakura_step(int adaptive, int mode, size_t n, void* const* arrays, double* scratch,
                      double* control, double eta, int capacity,
                      unsigned long long* counters, int* status, void* stream)
{
    if (mode != TUPAN_MODE_FP64 && mode != TUPAN_MODE_SPLIT32 && mode != TUPAN_MODE_IEEE64)
        return set_error(TUPAN_ERR_ARG, "unknown mode");
    if (capacity < 0) return set_error(TUPAN_ERR_ARG, "negative near-pair capacity");
    return sakura_step_call(adaptive, mode, n, arrays, scratch, control, eta, capacity,
                            counters, status, stream);
}

int tupan_sakura_graph(size_t ni, const double* const* ifields, size_t nj,
                       const double* const* jfields, double dt, long flag, double* const* out,
                       int capacity, unsigned long long* counters, int* status, void* stream)
{
    if (!ifields || !jfields || !out) return set_error(TUPAN_ERR_ARG, "null field list");
    if (capacity < 0) return set_error(TUPAN_ERR_ARG, "negative near-pair capacity");
    SakuraOptions opt;
    opt.path = 2;
    opt.graph_capacity = capacity;
    return sakura_call_ex(ni, ifields, nj, jfields, dt, flag, out, opt, counters, status,
                          stream);
}

int tupan_sia_subtree(int n, void* const* arrays, int narrays, const int* roles, int acc_mode,
                      int tstep_mode, double eta, double tau, int nbridge, const int* bridge_op,
                      const double* bridge_c, int nevolve, const int* evolve_op,
                      const double* evolve_c, double* t_curr, unsigned long long* pairs,
                      int* status, void* stream)
{
    return sia_subtree_call(n, arrays, narrays, roles, acc_mode, tstep_mode, eta, tau, nbridge,
                            bridge_op, bridge_c, nevolve, evolve_op, evolve_c, t_curr, pairs,
                            status, stream);
}

int tupan_kepler_pairs(size_t npairs, const int* idx, const double* m, double* rx, double* ry,
                       double* rz, const double* e2, double* vx, double* vy, double* vz,
                       double dt, int* status, void* stream)
{
    return kepler_pairs_call(npairs, idx, m, rx, ry, rz, e2, vx, vy, vz, dt, status, stream);
}

// ---- ensembles of independent systems --------------------------------------

int tupan_ensemble_create(int nsys, const int* offsets, int min_lanes, tupan_ensemble** plan)
{
    void* p = nullptr;
    int rc = ensemble_plan_create(nsys, offsets, min_lanes, &p);
    if (plan) *plan = static_cast<tupan_ensemble*>(p);
    return rc;
}

int tupan_ensemble_destroy(tupan_ensemble* plan) { return ensemble_plan_destroy(plan); }

int tupan_ensemble_info(const tupan_ensemble* plan, int* nsys, int* ntot, int* nwaves)
{
    return ensemble_plan_info(plan, nsys, ntot, nwaves);
}

int tupan_ensemble_set_guess(tupan_ensemble* plan, const void* net, unsigned long long* fallbacks)
{
    return ensemble_plan_set_guess(plan, net, fallbacks);
}

int tupan_ensemble_sakura(const tupan_ensemble* plan, const double* m, const double* rx,
                          const double* ry, const double* rz, const double* e2,
                          const double* vx, const double* vy, const double* vz, double dt,
                          long flag, double* drx, double* dry, double* drz, double* dvx,
                          double* dvy, double* dvz, unsigned long long* counters, int* status,
                          void* stream)
{
    const double* in[8] = {m, rx, ry, rz, e2, vx, vy, vz};
    double* out[6] = {drx, dry, drz, dvx, dvy, dvz};
    return ensemble_sakura_call(plan, in, dt, flag, out, counters, status, stream);
}

int tupan_ensemble_phi(const tupan_ensemble* plan, const double* m, const double* rx,
                       const double* ry, const double* rz, const double* e2, double* phi,
                       void* stream)
{
    const double* in[8] = {m, rx, ry, rz, e2, rx, ry, rz};
    return ensemble_phi_call(plan, in, phi, stream);
}

}  // extern "C"
