/*
 * tupan_hip.h -- C ABI of libtupan_hip.so, the gfx950 (MI355X) implementation
 * of tupan's all-pairs gravity kernels and Sakura two-body drift.
 *
 * Two layers:
 *
 * 1. Host-pointer layer: the exact names, argument orders and types of the
 *    reference's FP64 C backend (tupan/lib/src/libtupan.h:2-246 built with
 *    CONFIG_USE_DOUBLE: UINT = unsigned long, INT = long, REAL = double).
 *    Arrays are host memory; each call copies its inputs to the GPU, runs
 *    the FP64 kernel and copies the outputs back before returning, exactly
 *    like the reference's synchronous cffi calls.  libtupan_hip_f32.so
 *    exports the same names for the float build (UINT = unsigned int,
 *    INT = int, REAL = float), as the reference builds one library per
 *    precision (tupan/lib/cffi_backend.py:84-87).  These functions return
 *    void like the reference; on failure they fill every output with NaN,
 *    print the reason to stderr and leave it in tupan_last_error().
 *
 * 2. Device-resident layer (tupan_*): the same kernels on DEVICE pointers,
 *    enqueued on a caller-supplied HIP stream (NULL = the null stream), so
 *    particle arrays stay in HBM across steps.  Each returns TUPAN_OK or a
 *    negative status; tupan_last_error() gives the message.  `mode` selects
 *    the arithmetic of the FP64 entry points (TUPAN_MODE_*).
 *
 * Output arrays are overwritten, never accumulated into.  Kernels keep no
 * pointer past the call.  The device layer uses a per-device scratch
 * workspace for j-split partial sums (and, for Acc-Jerk calls of at least
 * 2^28 pairs in the fp64 and split32 modes, the hash table of a scan for
 * particles that share a position: it decides whether the pair loop may
 * leave out the test that drops zero-separation pairs; the sums are the
 * same either way, but for split32 pairs closer than half an FP32 ulp of
 * their softening, which then count as in the reference), so calls on one
 * device must not
 * overlap in time on different streams.  A workspace that a larger call
 * outgrows is replaced by one at least twice its size and kept allocated
 * until the process ends, so HIP graphs captured earlier stay valid.
 */
#ifndef TUPAN_HIP_H
#define TUPAN_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define TUPAN_OK 0
#define TUPAN_ERR_ARG (-1)     /* bad argument (null pointer, size, flag) */
#define TUPAN_ERR_HIP (-2)     /* HIP runtime error (launch, copy, device) */
#define TUPAN_ERR_ALLOC (-3)   /* device allocation failed */
#define TUPAN_ERR_SOLVER (-4)  /* Kepler solver hit its iteration bound */

/* Arithmetic modes of the FP64 device entry points. */
#define TUPAN_MODE_FP64 0     /* FP64, refined hardware rsqrt, fused ops      */
#define TUPAN_MODE_SPLIT32 1  /* split-FP32 pair arithmetic, FP64 sums (Acc,
                                 AccJerk, Tstep, PNAcc; Phi, SnapCrackle and
                                 NREG fall back to FP64)                     */
#define TUPAN_MODE_IEEE64 2   /* FP64 in the reference's operation order with
                                 IEEE division and square root               */

const char *tupan_last_error(void);
int tupan_version(void);
/* Number of compute units of the current device, and the workgroup count the
 * launch planner aims for; both -1 if no device. */
int tupan_device_info(int *cu_count, int *target_workgroups);

/* Define TUPAN_HIP_NO_FP64_REFERENCE_ABI before including this header to
 * hide the FP64 reference-ABI names (the float build re-declares them with
 * float types). */
#ifndef TUPAN_HIP_NO_FP64_REFERENCE_ABI
/* Kernel timing (no counterpart in the reference; used by the benchmark).
 * tupan_timing(1) clears the record and brackets the main all-pairs kernel
 * of every later call with HIP events on the stream it is launched on;
 * tupan_timing_read waits for the recorded launches and returns the summed
 * kernel time in milliseconds and the number of launches.  Timing state is
 * process-wide and not thread-safe. */
int tupan_timing(int enable);
int tupan_timing_read(double *total_ms, int *count);
/* The recorded launches one by one: ms[k] for k < min(max, *count), in
 * launch order (no counterpart in the reference).  The two-pass Sakura form
 * records two intervals per call: the far-pair kernel, then the rest of
 * the call. */
int tupan_timing_read_each(double *ms, int max, int *count);

/* Launch tuning of the pairwise kernels (no counterpart in the reference,
 * whose OpenCL build fixes its tile with -D LSIZE/WPT/UNROLL macros,
 * opencl_backend.py:78-88).  Every kernel is compiled for TUPAN_TUNE_SHAPES
 * tile shapes; tupan_tune_shape describes shape s (threads per workgroup,
 * i-particles per thread, j-slots per LDS tile, j per unrolled step).  The
 * table entry for (kernel TUPAN_K_*, slot, bucket) picks the shape and the
 * workgroups-per-CU target of the launch planner.  slot: 0 FP64, 1
 * split-FP32, 2 reference-order FP64, 3 float build.  bucket by pair count
 * ni*nj: 0 below 2^22, 1 below 2^30, 2 above.  Default: shape 0, 4
 * workgroups per CU.  The table is process-wide.  Calls of at most 2^21
 * pairs with 256 < nj <= 2048 (any mode but reference-order FP64) run a
 * one-launch small-set kernel that does not consult the table. */
#define TUPAN_TUNE_SHAPES 6
#define TUPAN_TUNE_BUCKETS 3
int tupan_tune_set(int kernel, int slot, int bucket, int shape, int wg_per_cu);
int tupan_tune_get(int kernel, int slot, int bucket, int *shape,
                   int *wg_per_cu);
int tupan_tune_shape(int kernel, int shape, int *block, int *wpt, int *tile,
                     int *unroll);

/* Calibration of the split-FP32 1/R (no counterpart in the reference).  The
 * split kernels take 1/R from the hardware v_rsq_f32 estimate and divide
 * their FP32 sums by 1 + 3 b, b the estimate's mean relative error, which is
 * compiled in.  This measures b on the current device over every float in
 * [1, 4) (weight 1/x): out[0], out[1] the raw estimate's mean error on
 * [1, 2) and [2, 4), out[2], out[3] the same after one Newton step, out[4]
 * the compiled-in b (used for both binades).  Allocates 1.5 MiB of device
 * memory for the call. */
int tupan_rsq_bias(double out[5]);

/* ---------------------------------------------------------------------------
 * Host-pointer layer (reference ABI, FP64).  libtupan.h line numbers cite
 * the declaration each function replaces.
 * ------------------------------------------------------------------------- */

/* libtupan.h:2-15 */
void phi_kernel(const unsigned long ni,
                const double *_im, const double *_irx, const double *_iry,
                const double *_irz, const double *_ie2,
                const unsigned long nj,
                const double *_jm, const double *_jrx, const double *_jry,
                const double *_jrz, const double *_je2,
                double *_iphi);

/* libtupan.h:17-32 */
void acc_kernel(const unsigned long ni,
                const double *_im, const double *_irx, const double *_iry,
                const double *_irz, const double *_ie2,
                const unsigned long nj,
                const double *_jm, const double *_jrx, const double *_jry,
                const double *_jrz, const double *_je2,
                double *_iax, double *_iay, double *_iaz);

/* libtupan.h:34-58 */
void acc_jerk_kernel(const unsigned long ni,
                     const double *_im, const double *_irx, const double *_iry,
                     const double *_irz, const double *_ie2,
                     const double *_ivx, const double *_ivy, const double *_ivz,
                     const unsigned long nj,
                     const double *_jm, const double *_jrx, const double *_jry,
                     const double *_jrz, const double *_je2,
                     const double *_jvx, const double *_jvy, const double *_jvz,
                     double *_iax, double *_iay, double *_iaz,
                     double *_ijx, double *_ijy, double *_ijz);

/* libtupan.h:60-96 */
void snap_crackle_kernel(const unsigned long ni,
                         const double *_im, const double *_irx,
                         const double *_iry, const double *_irz,
                         const double *_ie2, const double *_ivx,
                         const double *_ivy, const double *_ivz,
                         const double *_iax, const double *_iay,
                         const double *_iaz, const double *_ijx,
                         const double *_ijy, const double *_ijz,
                         const unsigned long nj,
                         const double *_jm, const double *_jrx,
                         const double *_jry, const double *_jrz,
                         const double *_je2, const double *_jvx,
                         const double *_jvy, const double *_jvz,
                         const double *_jax, const double *_jay,
                         const double *_jaz, const double *_jjx,
                         const double *_jjy, const double *_jjz,
                         double *_isx, double *_isy, double *_isz,
                         double *_icx, double *_icy, double *_icz);

/* libtupan.h:98-119 */
void tstep_kernel(const unsigned long ni,
                  const double *_im, const double *_irx, const double *_iry,
                  const double *_irz, const double *_ie2,
                  const double *_ivx, const double *_ivy, const double *_ivz,
                  const unsigned long nj,
                  const double *_jm, const double *_jrx, const double *_jry,
                  const double *_jrz, const double *_je2,
                  const double *_jvx, const double *_jvy, const double *_jvz,
                  const double eta,
                  double *_idt_a, double *_idt_b);

/* libtupan.h:152-178 */
void nreg_Xkernel(const unsigned long ni,
                  const double *_im, const double *_irx, const double *_iry,
                  const double *_irz, const double *_ie2,
                  const double *_ivx, const double *_ivy, const double *_ivz,
                  const unsigned long nj,
                  const double *_jm, const double *_jrx, const double *_jry,
                  const double *_jrz, const double *_je2,
                  const double *_jvx, const double *_jvy, const double *_jvz,
                  const double dt,
                  double *_idrx, double *_idry, double *_idrz,
                  double *_iax, double *_iay, double *_iaz, double *_iu);

/* libtupan.h:180-201 */
void nreg_Vkernel(const unsigned long ni,
                  const double *_im, const double *_ivx, const double *_ivy,
                  const double *_ivz, const double *_iax, const double *_iay,
                  const double *_iaz,
                  const unsigned long nj,
                  const double *_jm, const double *_jvx, const double *_jvy,
                  const double *_jvz, const double *_jax, const double *_jay,
                  const double *_jaz,
                  const double dt,
                  double *_idvx, double *_idvy, double *_idvz, double *_ik);

/* libtupan.h:121-149.  Post-Newtonian pair accelerations: order 2 (or 3)
 * gives 1PN, 4 adds 2PN, 5 adds 2.5PN, 6 adds 3PN, 7 (or more) adds 3.5PN;
 * order 0 or 1 gives zeros.  inv1..inv7 are 1/c^k (inv1 and inv3 are not
 * read, as in the reference). */
void pnacc_kernel(const unsigned long ni,
                  const double *_im, const double *_irx, const double *_iry,
                  const double *_irz, const double *_ie2,
                  const double *_ivx, const double *_ivy, const double *_ivz,
                  const unsigned long nj,
                  const double *_jm, const double *_jrx, const double *_jry,
                  const double *_jrz, const double *_je2,
                  const double *_jvx, const double *_jvy, const double *_jvz,
                  const unsigned long order, const double inv1,
                  const double inv2, const double inv3, const double inv4,
                  const double inv5, const double inv6, const double inv7,
                  double *_ipnax, double *_ipnay, double *_ipnaz);

/* libtupan.h:203-229 */
void sakura_kernel(const unsigned long ni,
                   const double *_im, const double *_irx, const double *_iry,
                   const double *_irz, const double *_ie2,
                   const double *_ivx, const double *_ivy, const double *_ivz,
                   const unsigned long nj,
                   const double *_jm, const double *_jrx, const double *_jry,
                   const double *_jrz, const double *_je2,
                   const double *_jvx, const double *_jvy, const double *_jvz,
                   const double dt, const long flag,
                   double *_idrx, double *_idry, double *_idrz,
                   double *_idvx, double *_idvy, double *_idvz);

/* libtupan.h:231-246.  Two particles (index 0 and 1); outputs may alias the
 * inputs, as the reference extension layer passes them (extensions.py:
 * 642-646). */
void kepler_solver_kernel(const double *_im, const double *_irx,
                          const double *_iry, const double *_irz,
                          const double *_ie2, const double *_ivx,
                          const double *_ivy, const double *_ivz,
                          const double dt,
                          double *_ir1x, double *_ir1y, double *_ir1z,
                          double *_iv1x, double *_iv1y, double *_iv1z);
#endif /* TUPAN_HIP_NO_FP64_REFERENCE_ABI */

/* ---------------------------------------------------------------------------
 * Device-resident layer (FP64 arrays).  Same argument order as above, with
 * size_t counts, a mode and a trailing stream.  Replaces the same reference
 * functions (libtupan.h lines as above) for callers that keep particles in
 * HBM.
 * ------------------------------------------------------------------------- */
int tupan_phi(int mode, size_t ni,
              const double *im, const double *irx, const double *iry,
              const double *irz, const double *ie2,
              size_t nj,
              const double *jm, const double *jrx, const double *jry,
              const double *jrz, const double *je2,
              double *iphi, void *stream);

int tupan_acc(int mode, size_t ni,
              const double *im, const double *irx, const double *iry,
              const double *irz, const double *ie2,
              size_t nj,
              const double *jm, const double *jrx, const double *jry,
              const double *jrz, const double *je2,
              double *iax, double *iay, double *iaz, void *stream);

int tupan_acc_jerk(int mode, size_t ni,
                   const double *im, const double *irx, const double *iry,
                   const double *irz, const double *ie2,
                   const double *ivx, const double *ivy, const double *ivz,
                   size_t nj,
                   const double *jm, const double *jrx, const double *jry,
                   const double *jrz, const double *je2,
                   const double *jvx, const double *jvy, const double *jvz,
                   double *iax, double *iay, double *iaz,
                   double *ijx, double *ijy, double *ijz, void *stream);

int tupan_snap_crackle(int mode, size_t ni,
                       const double *const *ifields /* 14: m rx ry rz e2 vx
                                                      vy vz ax ay az jx jy jz */,
                       size_t nj, const double *const *jfields /* same 14 */,
                       double *isx, double *isy, double *isz,
                       double *icx, double *icy, double *icz, void *stream);

int tupan_tstep(int mode, size_t ni,
                const double *im, const double *irx, const double *iry,
                const double *irz, const double *ie2,
                const double *ivx, const double *ivy, const double *ivz,
                size_t nj,
                const double *jm, const double *jrx, const double *jry,
                const double *jrz, const double *je2,
                const double *jvx, const double *jvy, const double *jvz,
                double eta, double *idt_a, double *idt_b, void *stream);

/* Generic form of the five pairwise kernels above: `kernel` is one of
 * TUPAN_K_*; ifields/jfields hold the reference's input arrays in its order
 * (Phi, Acc: m rx ry rz e2; AccJerk, Tstep: + vx vy vz; SnapCrackle: + ax ay
 * az jx jy jz); out holds the outputs in the reference's order; eta is read
 * by Tstep only. */
#define TUPAN_K_PHI 0
#define TUPAN_K_ACC 1
#define TUPAN_K_ACC_JERK 2
#define TUPAN_K_SNAP_CRACKLE 3
#define TUPAN_K_TSTEP 4
#define TUPAN_K_NREG_X 5  /* nreg_Xkernel: 8 + 8 fields, dt; 7 outputs */
#define TUPAN_K_NREG_V 6  /* nreg_Vkernel: 7 + 7 fields, dt; 4 outputs */
#define TUPAN_K_PNACC 7   /* pnacc_kernel: 8 + 8 fields; through tupan_pnacc */
int tupan_pairwise(int kernel, int mode, size_t ni,
                   const double *const *ifields, size_t nj,
                   const double *const *jfields, double *const *out,
                   double eta, void *stream);
/* FP32 arrays (the reference's float build); mode is ignored. */
int tupan_pairwise_f32(int kernel, int mode, size_t ni,
                       const float *const *ifields, size_t nj,
                       const float *const *jfields, float *const *out,
                       float eta, void *stream);

/* Frees the scratch workspaces of the current device after a device
 * synchronisation: the current ones and those retired when they grew (no
 * counterpart in the reference, whose C backend allocates nothing on a
 * device).  The library's pairwise calls keep one growable workspace per
 * device; an outgrown buffer stays allocated until this call because HIP
 * graphs captured earlier still address it, so graphs that contain library
 * launches must be captured again after it.  The next call allocates
 * afresh.  Calls on different streams that use the same workspace are
 * ordered on the device (the later stream waits for the earlier one's work,
 * so the earlier stream must not have been destroyed) unless a stream is
 * capturing a graph: a capture must not overlap other streams' calls. */
int tupan_release_workspaces(void);

/* tupan_pairwise without allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as tupan_pairwise, except that the j-split
 * partial sums go to a fixed device array of TUPAN_NOALLOC_SCRATCH_BYTES
 * (64 MiB) that the library holds on every device it runs on until the
 * process ends, instead of the growable workspace, and that no timing
 * events are recorded.  The launch plan caps its j-split count at what that
 * array holds (splits x partial fields x ni doubles), so a call that needs
 * more partial sums than the maximum runs with fewer j-splits (down to one,
 * which needs none) rather than overflowing it: results equal
 * tupan_pairwise's bit for bit when the cap does not bind, and to FP64
 * summation order otherwise.  A plan that still exceeded the array would
 * return TUPAN_ERR_ARG.  The first call on a device looks up that device's
 * array (hipGetSymbolAddress) and fails with TUPAN_ERR_HIP if it cannot.
 * Every call on one device shares that device's array: calls that overlap
 * in time on one device (different streams, threads, or replays of graphs
 * that hold this call) race on it and give undefined results.  The array
 * needs no clearing between calls: every partial a call combines is
 * written by that call first.  The coincidence scan of large Acc-Jerk
 * calls is left out, so their tiles keep the zero-separation test: the same
 * sums, but for split32 pairs closer than half an FP32 ulp of their
 * softening, which tupan_pairwise then counts.  TUPAN_K_PNACC returns
 * TUPAN_ERR_ARG, as for tupan_pairwise. */
#define TUPAN_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int tupan_pairwise_noalloc(int kernel, int mode, size_t ni,
                           const double *const *ifields, size_t nj,
                           const double *const *jfields, double *const *out,
                           double eta, void *stream);

/* PN pair accelerations (pnacc_kernel.c:5-58, libtupan.h:121-149) on device
 * arrays.  ifields/jfields: m rx ry rz e2 vx vy vz; out: pnax pnay pnaz;
 * inv7[k-1] = 1/c^k for k = 1..7 (a host array).  Modes TUPAN_MODE_FP64 and
 * TUPAN_MODE_IEEE64 (bit-exact with the reference C backend for one j
 * split) and TUPAN_MODE_SPLIT32 (pair terms in FP32 on hi+lo separations,
 * FP64 sums; about 1e-6 of the PN acceleration, which is itself (v/c)^2 of
 * the Newtonian one).  tupan_pairwise with TUPAN_K_PNACC
 * returns TUPAN_ERR_ARG: the PN parameters only travel through these. */
int tupan_pnacc(int mode, size_t ni, const double *const *ifields, size_t nj,
                const double *const *jfields, unsigned long order,
                const double *inv7, double *const *out, void *stream);
int tupan_pnacc_f32(int mode, size_t ni, const float *const *ifields,
                    size_t nj, const float *const *jfields,
                    unsigned long order, const float *inv7,
                    float *const *out, void *stream);

/* Sakura pair drift (sakura_kernel.c:5-63, libtupan.h:203-229) on device
 * arrays.  flag: -1, +1, -2, +2 (any other value leaves every pair
 * unchanged, as in the reference).  `counters`, if not NULL, is a device
 * array of 4 unsigned 64-bit integers that receives (added to, not
 * overwritten) the numbers of Kepler solves, Laguerre iterations, step
 * subdivisions and leapfrog pair steps; `status` if not NULL is a device int
 * set to a nonzero value if any pair exhausted the solver's bounded retries.
 */
int tupan_sakura(size_t ni,
                 const double *im, const double *irx, const double *iry,
                 const double *irz, const double *ie2,
                 const double *ivx, const double *ivy, const double *ivz,
                 size_t nj,
                 const double *jm, const double *jrx, const double *jry,
                 const double *jrz, const double *je2,
                 const double *jvx, const double *jvy, const double *jvz,
                 double dt, long flag,
                 double *idrx, double *idry, double *idrz,
                 double *idvx, double *idvy, double *idvz,
                 unsigned long long *counters, int *status, void *stream);

/* Kepler two-body advance of many independent pairs (the SIA n = 2 leaves,
 * kepler_solver_kernel.c:5-49) in one launch: pair p uses particles
 * idx[2p] and idx[2p+1] of the device SoA arrays and advances them in place
 * by dt.  npairs may be 0.  status as for tupan_sakura. */
int tupan_kepler_pairs(size_t npairs, const int *idx,
                       const double *m, double *rx, double *ry, double *rz,
                       const double *e2, double *vx, double *vy, double *vz,
                       double dt, int *status, void *stream);

/* Sakura with the options of the two-pass form (no counterpart in the
 * reference; same arithmetic as tupan_sakura, sakura_kernel.c:5-63).
 * Fields are the 8 input arrays (m, rx, ry, rz, e2, vx, vy, vz) of each side
 * and the 6 outputs (drx, dry, drz, dvx, dvy, dvz), device pointers.
 *   path   0: automatic (two-pass for ni * nj >= 2^22), 1: one-pass,
 *          2: two-pass (far pairs in the tile loop, near pairs compacted
 *          into a dense list and solved one per thread)
 *   net    packed Kepler-guess network (tupan_mlp_pack; device, 16-byte
 *          aligned) or NULL: the near pass evaluates the network on each
 *          near pair's first solve, which starts from the network's guess
 *          and falls back to the plain guess after 8 Laguerre steps
 *   log    NULL or a device log of the near pairs' first solves
 *   counters  NULL or 5 device counters: solves, Laguerre steps,
 *          subdivisions, leapfrog pairs, guess fallbacks
 *   near_pairs  NULL or host int: near pairs listed (-1: one-pass form)
 * net or log imply the two-pass form.  One host synchronisation per call
 * (the near-pair count). */
typedef struct {
    float *features;  /* [4][capacity]: tau, sigma, kappa, eps (kepler.h) */
    float *target;    /* [capacity]: s r / dt - 1 of the converged first
                         solve, NaN if it did not converge undivided     */
    int *iterations;  /* [capacity]: Laguerre steps of the pair           */
    int capacity;
} tupan_solve_log;

int tupan_sakura_ex(size_t ni, const double *const *ifields, size_t nj,
                    const double *const *jfields, double dt, long flag,
                    double *const *out, int path, const void *net,
                    const tupan_solve_log *log, unsigned long long *counters,
                    int *near_pairs, int *status, void *stream);

/* Sakura pair drift for HIP graph capture (no counterpart in the
 * reference).  The two-pass form of tupan_sakura_ex without its one host
 * synchronisation: the near-pair list gets `capacity` entries reserved, the
 * near passes read the listed count on the device, and a call with more near
 * pairs than `capacity` sets bit 2 of *status (required, device int) and
 * leaves its outputs invalid, so the caller checks status after each replay
 * and retries with a larger capacity.  Make one call with the same sizes and
 * capacity outside capture first (it sizes the library's workspaces; a
 * capture must not grow them).  counters as for tupan_sakura. */
int tupan_sakura_graph(size_t ni, const double *const *ifields, size_t nj,
                       const double *const *jfields, double dt, long flag,
                       double *const *out, int capacity,
                       unsigned long long *counters, int *status, void *stream);

/* Hierarchical SIA step of a small subset on the device (no counterpart in
 * the reference, whose driver visits every level from the host:
 * tupan/integrator/sia.py:1089-1123, 'h' variants).  Runs recurse(S, tau)
 * for the n <= 64 particles of S in one launch of one wave, with the host
 * driver's operations in the host driver's order, so the results are the
 * host driver's bit for bit.
 *   arrays[narrays]  every 8-byte per-particle array of S (device; they are
 *                    permuted together at each level's [slow | fast] split)
 *   roles[15]        index in arrays of mass, eps2, rx, ry, rz, vx, vy, vz,
 *                    time, nstep (int64), tstep, tstepij, ax, ay, az
 *   acc_mode, tstep_mode  TUPAN_MODE_FP64 or TUPAN_MODE_SPLIT32
 *   bridge_op/c[nbridge]  the scheme's slow/fast sequence: 0 = evolve slow
 *                    and recurse on fast over c tau, 1 = slow<->fast kicks
 *   evolve_op/c[nevolve]  the scheme's sequence on a set: 0 = drift,
 *                    1 = kick, over c dt
 *   t_curr           device double, advanced as the host driver advances
 *                    its time (by tau at each level whose fast set is empty)
 *   pairs            device uint64[2], incremented by the Tstep and Acc
 *                    interactions evaluated
 *   status           device int: 1 Kepler failure, 2 more than 64 levels,
 *                    3 more than 65536 level visits */
int tupan_sia_subtree(int n, void *const *arrays, int narrays, const int *roles,
                      int acc_mode, int tstep_mode, double eta, double tau,
                      int nbridge, const int *bridge_op, const double *bridge_c,
                      int nevolve, const int *evolve_op, const double *evolve_c,
                      double *t_curr, unsigned long long *pairs, int *status,
                      void *stream);

/* Device-driven shared-step SIA steps: the 's' (fixed step) and 'a'
 * (adaptive block step) variants of the SIA methods
 * (tupan/integrator/sia.py:1089-1123 with the block step of
 * integrator/__init__.py:48-78), in which every particle is slow at every
 * step.  No counterpart in the reference (its driver loop is Python); the
 * arithmetic is the reference driver's.  One call enqueues a whole step on
 * `stream`: for adaptive steps Tstep (tstep_mode, eta) into scratch, the step
 * size into the control block as tupan_hermite_step computes it, then for
 * k < nops: ops[k] = 0 drift r += v h, 1 kick (Acc(all, all) in acc_mode
 * into scratch, copied to the a arrays, v += a h), with
 * h = coef_evolve[k] * (coef_bridge[k] * tau) (the scheme's bridge and
 * evolve coefficients, nops <= 1024), then tstepij (adaptive), tstep = tau,
 * time += tau, nstep += 1, t_curr += tau.
 *   arrays[TUPAN_SIA_*]  device arrays of n >= 3 values: FP64, except nstep
 *                        (int64)
 *   scratch              device FP64 workspace of 5 n values
 *   control              device FP64 [5]: t_curr, t_end, eta, tau, status, as
 *                        for tupan_hermite_step (a step at or past t_end
 *                        changes nothing) */
#define TUPAN_SIA_MASS 0
#define TUPAN_SIA_EPS2 1
#define TUPAN_SIA_RX 2     /* RX RY RZ */
#define TUPAN_SIA_VX 5     /* VX VY VZ */
#define TUPAN_SIA_AX 8     /* AX AY AZ */
#define TUPAN_SIA_TSTEP 11
#define TUPAN_SIA_TSTEPIJ 12
#define TUPAN_SIA_TIME 13
#define TUPAN_SIA_NSTEP 14
#define TUPAN_SIA_NROLES 15
int tupan_sia_step(int adaptive, int acc_mode, int tstep_mode, size_t n,
                   void *const *arrays, double *scratch, double *control,
                   double eta, int nops, const int *ops,
                   const double *coef_bridge, const double *coef_evolve,
                   void *stream);

/* Device-scheduled hierarchical SIA steps: the 'h' variants of the SIA
 * methods (tupan/integrator/sia.py:1089-1123) with every level visit decided
 * on the device.  No counterpart in the reference, whose driver visits the
 * levels from Python; the operations, their order and every pairwise call's
 * launch plan are the host driver's, so the results are its bit for bit.
 * A plan holds the walk's state in device memory (a stack of level frames)
 * and is tied to the particle arrays, the scheme and the tuning table at
 * creation.  tupan_sia_levels_enqueue enqueues `nslots` slots on `stream`
 * (with begin != 0 first the step's start: the base step from the control
 * time, t_end and eta, and the walk at the whole set); each slot performs
 * one operation of the walk -- a Tstep or Acc call, a [slow | fast]
 * partition, a drift, kick or bookkeeping update, a subset of at most 64
 * particles in one launch (tupan_sia_subtree), a one- or two-particle evolve
 * -- as a fixed sequence of launches whose workers read the operation from
 * device memory, so any run of slots can be captured as one HIP graph.  The
 * step is complete when the state's done flag is set (slots after that do
 * nothing); a step takes roughly 12 slots per level visit above 64
 * particles.
 *   arrays[narrays]   every 8-byte per-particle array of the system (device;
 *                     permuted together by the partitions), narrays <= 32
 *   roles[15], acc_mode, tstep_mode, eta, bridge/evolve sequences: as for
 *                     tupan_sia_subtree; n > 64
 *   set:   control time t_curr, t_end and eta (synchronous); clear_pairs
 *          zeroes the Tstep and Acc interaction counters
 *   state: step[5] = t_curr, t_end, eta, tau, status of the step size;
 *          state[4] = done, status (1 Kepler failure, 2 more than 64 levels,
 *          3 more than 2^20 visits, 4 operation queue overflow, 16 scheduler
 *          bound, 32 no valid step size), level visits above 64 particles,
 *          slots that did work; pairs[2] = Tstep and Acc interactions
 *          (synchronous) */
typedef struct tupan_sia_levels tupan_sia_levels;
int tupan_sia_levels_create(int n, void *const *arrays, int narrays,
                            const int *roles, int acc_mode, int tstep_mode,
                            double eta, int nbridge, const int *bridge_op,
                            const double *bridge_c, int nevolve,
                            const int *evolve_op, const double *evolve_c,
                            tupan_sia_levels **plan);
int tupan_sia_levels_destroy(tupan_sia_levels *plan);
int tupan_sia_levels_set(tupan_sia_levels *plan, double t_curr, double t_end,
                         double eta, int clear_pairs);
int tupan_sia_levels_enqueue(tupan_sia_levels *plan, int begin, int nslots,
                             void *stream);
int tupan_sia_levels_state(tupan_sia_levels *plan, double step[5],
                           int state[4], unsigned long long pairs[2]);

/* ---------------------------------------------------------------------------
 * Kepler-guess network (no counterpart in the reference): 4 -> 64 -> 64 -> 1
 * perceptron with SiLU, inputs x = (log tau, sigma, kappa, eps) of
 * tupan_solve_log standardised by norm[0:4] (mean) and norm[4:8] (1/std),
 * output y mapped to delta = y * norm[8] + norm[9].  Parameters: FP32, flat
 * W1[64][4] b1[64] W2[64][64] b2[64] W3[64] b3.  All arrays device memory.
 * ------------------------------------------------------------------------- */
int tupan_mlp_sizes(int *nparams, int *nnorm, int *packed_bytes);
/* Mean squared error of y against (target - norm[9]) / norm[8] over n
 * samples and its gradient (FP32, deterministic). */
int tupan_mlp_grad(int n, const float *features, const float *target,
                   const float *params, const float *norm, float *grad,
                   float *loss, void *stream);
/* FP32 forward: delta for n samples. */
int tupan_mlp_forward(int n, const float *features, const float *params,
                      const float *norm, float *delta, void *stream);
/* One Adam step (step counts from 1). */
int tupan_mlp_adam(float *params, const float *grad, float *m1, float *m2,
                   float lr, float beta1, float beta2, float eps, int step,
                   void *stream);
/* BF16 inference block (tupan_mlp_sizes: packed_bytes) from FP32 params. */
int tupan_mlp_pack(const float *params, const float *norm, void *packed,
                   void *stream);
/* BF16 inference with FP32 accumulation (matrix cores for the 64x64 layer). */
int tupan_mlp_infer(int n, const float *features, const void *packed,
                    float *delta, void *stream);

/* ---------------------------------------------------------------------------
 * Ensembles of independent few-body systems (no counterpart in the reference,
 * which integrates one system per process; the per-pair arithmetic is
 * sakura_kernel.c:5-63 and phi_kernel.c:5-34 restricted to pairs inside a
 * system).  The particles of all systems are concatenated in one SoA set;
 * system s is [offsets[s], offsets[s+1]), each of at most 64 particles.
 * A plan (created once from HOST offsets) packs systems of similar size into
 * wavefronts, min_lanes (8, 16, 32 or 64) lanes per system at least; every
 * call on the plan is one kernel launch over all systems.
 * ------------------------------------------------------------------------- */
typedef struct tupan_ensemble tupan_ensemble;
int tupan_ensemble_create(int nsys, const int *offsets, int min_lanes,
                          tupan_ensemble **plan);
int tupan_ensemble_destroy(tupan_ensemble *plan);
int tupan_ensemble_info(const tupan_ensemble *plan, int *nsys, int *ntot,
                        int *nwaves);
/* Kepler-guess network for the plan's Sakura calls (no counterpart in the
 * reference; the learned guess of tupan_sakura_ex applied to ensembles).
 * net: packed network (tupan_mlp_pack; device, 16-byte aligned) or NULL for
 * the plain guess.  With a network every j step of a wave evaluates it on
 * the wave's 64 pairs and each pair's first solve starts from its guess
 * (falling back to the plain guess after 8 Laguerre steps); converged
 * states equal the plain solver's to its tolerance.  fallbacks: NULL or one
 * device counter the calls add their guess fallbacks to.  Both pointers are
 * kept by the plan and must outlive its calls. */
int tupan_ensemble_set_guess(tupan_ensemble *plan, const void *net,
                             unsigned long long *fallbacks);
/* Sakura pair drift inside every system (flag and counters/status as for
 * tupan_sakura). */
int tupan_ensemble_sakura(const tupan_ensemble *plan, const double *m,
                          const double *rx, const double *ry, const double *rz,
                          const double *e2, const double *vx, const double *vy,
                          const double *vz, double dt, long flag,
                          double *drx, double *dry, double *drz,
                          double *dvx, double *dvy, double *dvz,
                          unsigned long long *counters, int *status,
                          void *stream);
/* Potential of every particle due to its own system (reference operation
 * order, IEEE division and square root). */
int tupan_ensemble_phi(const tupan_ensemble *plan, const double *m,
                       const double *rx, const double *ry, const double *rz,
                       const double *e2, double *phi, void *stream);

/* ---------------------------------------------------------------------------
 * Device-driven Hermite steps (hermite.py:20-410 with the shared block step
 * of integrator/__init__.py:48-78).  One call enqueues one whole P(EC)^2
 * step of order 2, 4, 6 or 8 ('adaptive' = the 'a' methods: Tstep and the
 * power-of-two block step) on `stream`; the step size is computed on the
 * device from the control block, so a sequence of calls can be captured as
 * one HIP graph.  No counterpart in the reference (its driver loop is
 * Python); the arithmetic is the reference driver's.
 *
 *   arrays[TUPAN_H_*]  device arrays of n values: FP64, except nstep (int64);
 *                      s and c only for orders 6 and 8
 *   scratch            device FP64 workspace of 32 n values
 *   control            device FP64 [5]: t_curr, t_end, eta, tau, status; the
 *                      call reads t_curr, t_end and eta, writes tau and
 *                      advances t_curr; once t_curr reaches t_end a call
 *                      changes nothing (tau = 0); status (caller-zeroed)
 *                      becomes 1 for a non-finite t_curr, t_end or eta, 2
 *                      when Tstep gives no positive finite step, 3 when no
 *                      aligned block step exists, and then every later call
 *                      changes nothing; the host must check it
 *   eta                the same eta as control[2], for the Tstep kernel
 *   mode               force mode (TUPAN_MODE_*) of Acc/AccJerk/Tstep
 * ------------------------------------------------------------------------- */
#define TUPAN_H_MASS 0
#define TUPAN_H_EPS2 1
#define TUPAN_H_RX 2   /* RX RY RZ */
#define TUPAN_H_VX 5   /* VX VY VZ */
#define TUPAN_H_AX 8   /* AX AY AZ */
#define TUPAN_H_JX 11  /* JX JY JZ */
#define TUPAN_H_SX 14  /* SX SY SZ */
#define TUPAN_H_CX 17  /* CX CY CZ */
#define TUPAN_H_TSTEP 20
#define TUPAN_H_TSTEPIJ 21
#define TUPAN_H_TIME 22
#define TUPAN_H_NSTEP 23
#define TUPAN_H_NROLES 24
int tupan_hermite_step(int order, int adaptive, int mode, size_t n,
                       void *const *arrays, double *scratch, double *control,
                       double eta, void *stream);

/* ---------------------------------------------------------------------------
 * Device-driven Sakura steps (sakura.py:22-50; 'asakura' step size
 * sakura.py:100-117 with the block step of integrator/__init__.py:48-78).
 * One call enqueues one whole step: drift tau/2, pair drift (flag -1, tau/2)
 * and update, pair drift (flag +1) and update, drift tau/2, bookkeeping;
 * the pair drifts are the graph-safe two-pass form (tupan_sakura_graph) with
 * dt read from the control block, so a sequence of calls can be captured as
 * one HIP graph.  No counterpart in the reference (its driver loop is
 * Python); the arithmetic is the reference driver's.
 *
 *   arrays[TUPAN_S_*]  device arrays of n values: FP64, except nstep (int64);
 *                      DRX..DVZ receive each pair drift's summed changes
 *   scratch            device FP64 workspace of 2 n values ('asakura' only)
 *   control            device FP64 [6]: t_curr, t_end, eta, tau, status, dt;
 *                      as for tupan_hermite_step (status 2: no finite Tstep
 *                      maximum); dt = tau / 2 is what the pair drifts read
 *   capacity, counters, status  as for tupan_sakura_graph (status: device
 *                      int; bit 1 Kepler failure, bit 2 near-pair list
 *                      longer than capacity: the step's state is invalid)
 *   mode               force mode (TUPAN_MODE_*) of Tstep
 * ------------------------------------------------------------------------- */
#define TUPAN_S_MASS 0
#define TUPAN_S_EPS2 1
#define TUPAN_S_RX 2    /* RX RY RZ */
#define TUPAN_S_VX 5    /* VX VY VZ */
#define TUPAN_S_DRX 8   /* DRX DRY DRZ DVX DVY DVZ */
#define TUPAN_S_TSTEP 14
#define TUPAN_S_TSTEPIJ 15
#define TUPAN_S_TIME 16
#define TUPAN_S_NSTEP 17
#define TUPAN_S_NROLES 18
int tupan_sakura_step(int adaptive, int mode, size_t n, void *const *arrays,
                      double *scratch, double *control, double eta,
                      int capacity, unsigned long long *counters,
                      int *status, void *stream);

/* ---------------------------------------------------------------------------
 * Host layer of the float build.  libtupan_hip_f32.so implements the
 * reference's float ABI (phi_kernel, ... with unsigned int / int / float)
 * on top of these; ifields/jfields/out as for tupan_pairwise.  Sakura and
 * Kepler of the float build compute in FP64 and round to float.  Return
 * TUPAN_OK or a negative status.
 * ------------------------------------------------------------------------- */
int tupan_host_pairwise_f32(int kernel, size_t ni, const float *const *ifields,
                            size_t nj, const float *const *jfields,
                            float *const *out, float eta);
int tupan_host_pnacc_f32(size_t ni, const float *const *ifields, size_t nj,
                         const float *const *jfields, unsigned long order,
                         const float *inv7, float *const *out);
int tupan_host_sakura_f32(size_t ni, const float *const *ifields, size_t nj,
                          const float *const *jfields, float dt, int flag,
                          float *const *out);
/* in8: m rx ry rz e2 vx vy vz of two particles; out6: rx ry rz vx vy vz */
int tupan_host_kepler_f32(const float *const *in8, float dt,
                          float *const *out6);

#ifdef __cplusplus
}
#endif

#endif /* TUPAN_HIP_H */
