/*
 * CPU ORACLE, OpenMP variant -- TEST / CPU-BASELINE INFRASTRUCTURE ONLY.
 *
 * The same arithmetic as oracle.c (included below, so the per-pair code is
 * shared), with the outer i loop split across threads.  Each i row is still
 * summed sequentially over j, so the results are bitwise identical to the
 * serial oracle.  Used only to time the "all host cores" CPU baseline.
 */
#include "oracle.c"
#include <omp.h>

#ifdef ORACLE_SP
#define OSYM(name) oracle_omp_##name##_f32
#else
#define OSYM(name) oracle_omp_##name
#endif

/* Rows are handed out in contiguous chunks; each chunk calls the serial
 * kernel on its slice of i with the full j set. */
#define ROW_SPLIT(call_slice)                                               \
    do {                                                                    \
        _Pragma("omp parallel")                                             \
        {                                                                   \
            int nt = omp_get_num_threads(), t = omp_get_thread_num();       \
            size_t lo = ni * (size_t)t / (size_t)nt;                        \
            size_t hi = ni * (size_t)(t + 1) / (size_t)nt;                  \
            if (hi > lo) { size_t o = lo; size_t n = hi - lo; call_slice; } \
        }                                                                   \
    } while (0)

int OSYM(threads)(void) { return omp_get_max_threads(); }

void OSYM(phi)(size_t ni, const real *im, const real *irx, const real *iry,
               const real *irz, const real *ie2,
               size_t nj, const real *jm, const real *jrx, const real *jry,
               const real *jrz, const real *je2, real *iphi)
{
    ROW_SPLIT(SYM(phi)(n, im + o, irx + o, iry + o, irz + o, ie2 + o,
                       nj, jm, jrx, jry, jrz, je2, iphi + o));
}

void OSYM(acc)(size_t ni, const real *im, const real *irx, const real *iry,
               const real *irz, const real *ie2,
               size_t nj, const real *jm, const real *jrx, const real *jry,
               const real *jrz, const real *je2,
               real *iax, real *iay, real *iaz)
{
    ROW_SPLIT(SYM(acc)(n, im + o, irx + o, iry + o, irz + o, ie2 + o,
                       nj, jm, jrx, jry, jrz, je2, iax + o, iay + o, iaz + o));
}

void OSYM(acc_jerk)(size_t ni, const real *im, const real *irx,
                    const real *iry, const real *irz, const real *ie2,
                    const real *ivx, const real *ivy, const real *ivz,
                    size_t nj, const real *jm, const real *jrx,
                    const real *jry, const real *jrz, const real *je2,
                    const real *jvx, const real *jvy, const real *jvz,
                    real *iax, real *iay, real *iaz,
                    real *ijx, real *ijy, real *ijz)
{
    ROW_SPLIT(SYM(acc_jerk)(n, im + o, irx + o, iry + o, irz + o, ie2 + o,
                            ivx + o, ivy + o, ivz + o,
                            nj, jm, jrx, jry, jrz, je2, jvx, jvy, jvz,
                            iax + o, iay + o, iaz + o,
                            ijx + o, ijy + o, ijz + o));
}

void OSYM(tstep)(size_t ni, const real *im, const real *irx, const real *iry,
                 const real *irz, const real *ie2,
                 const real *ivx, const real *ivy, const real *ivz,
                 size_t nj, const real *jm, const real *jrx, const real *jry,
                 const real *jrz, const real *je2,
                 const real *jvx, const real *jvy, const real *jvz,
                 real eta, real *idt_a, real *idt_b)
{
    ROW_SPLIT(SYM(tstep)(n, im + o, irx + o, iry + o, irz + o, ie2 + o,
                         ivx + o, ivy + o, ivz + o,
                         nj, jm, jrx, jry, jrz, je2, jvx, jvy, jvz,
                         eta, idt_a + o, idt_b + o));
}

void OSYM(sakura)(size_t ni, const real *im, const real *irx, const real *iry,
                  const real *irz, const real *ie2,
                  const real *ivx, const real *ivy, const real *ivz,
                  size_t nj, const real *jm, const real *jrx, const real *jry,
                  const real *jrz, const real *je2,
                  const real *jvx, const real *jvy, const real *jvz,
                  real dt, sint flag,
                  real *idrx, real *idry, real *idrz,
                  real *idvx, real *idvy, real *idvz)
{
    ROW_SPLIT(SYM(sakura)(n, im + o, irx + o, iry + o, irz + o, ie2 + o,
                          ivx + o, ivy + o, ivz + o,
                          nj, jm, jrx, jry, jrz, je2, jvx, jvy, jvz,
                          dt, flag, idrx + o, idry + o, idrz + o,
                          idvx + o, idvy + o, idvz + o));
}
