// libtupan_hip_f32.so: the reference's float-build C ABI
// (tupan/lib/src/libtupan.h:2-246 without CONFIG_USE_DOUBLE: UINT = unsigned
// int, INT = int, REAL = float), forwarding to libtupan_hip.so.  The
// reference builds one library per precision with identical symbol names
// (tupan/lib/cffi_backend.py:84-87); this is the float one.
#include <stddef.h>

#define TUPAN_HIP_NO_FP64_REFERENCE_ABI
#include "tupan_hip.h"

extern "C" {

void phi_kernel(const unsigned int ni, const float* _im, const float* _irx, const float* _iry,
                const float* _irz, const float* _ie2, const unsigned int nj, const float* _jm,
                const float* _jrx, const float* _jry, const float* _jrz, const float* _je2,
                float* _iphi)
{
    const float* ih[5] = {_im, _irx, _iry, _irz, _ie2};
    const float* jh[5] = {_jm, _jrx, _jry, _jrz, _je2};
    float* oh[1] = {_iphi};
    tupan_host_pairwise_f32(TUPAN_K_PHI, ni, ih, nj, jh, oh, 0.0f);
}

void acc_kernel(const unsigned int ni, const float* _im, const float* _irx, const float* _iry,
                const float* _irz, const float* _ie2, const unsigned int nj, const float* _jm,
                const float* _jrx, const float* _jry, const float* _jrz, const float* _je2,
                float* _iax, float* _iay, float* _iaz)
{
    const float* ih[5] = {_im, _irx, _iry, _irz, _ie2};
    const float* jh[5] = {_jm, _jrx, _jry, _jrz, _je2};
    float* oh[3] = {_iax, _iay, _iaz};
    tupan_host_pairwise_f32(TUPAN_K_ACC, ni, ih, nj, jh, oh, 0.0f);
}

void acc_jerk_kernel(const unsigned int ni, const float* _im, const float* _irx,
                     const float* _iry, const float* _irz, const float* _ie2,
                     const float* _ivx, const float* _ivy, const float* _ivz,
                     const unsigned int nj, const float* _jm, const float* _jrx,
                     const float* _jry, const float* _jrz, const float* _je2,
                     const float* _jvx, const float* _jvy, const float* _jvz, float* _iax,
                     float* _iay, float* _iaz, float* _ijx, float* _ijy, float* _ijz)
{
    const float* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const float* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    float* oh[6] = {_iax, _iay, _iaz, _ijx, _ijy, _ijz};
    tupan_host_pairwise_f32(TUPAN_K_ACC_JERK, ni, ih, nj, jh, oh, 0.0f);
}

void snap_crackle_kernel(const unsigned int ni, const float* _im, const float* _irx,
                         const float* _iry, const float* _irz, const float* _ie2,
                         const float* _ivx, const float* _ivy, const float* _ivz,
                         const float* _iax, const float* _iay, const float* _iaz,
                         const float* _ijx, const float* _ijy, const float* _ijz,
                         const unsigned int nj, const float* _jm, const float* _jrx,
                         const float* _jry, const float* _jrz, const float* _je2,
                         const float* _jvx, const float* _jvy, const float* _jvz,
                         const float* _jax, const float* _jay, const float* _jaz,
                         const float* _jjx, const float* _jjy, const float* _jjz, float* _isx,
                         float* _isy, float* _isz, float* _icx, float* _icy, float* _icz)
{
    const float* ih[14] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy,
                           _ivz, _iax, _iay, _iaz, _ijx, _ijy, _ijz};
    const float* jh[14] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy,
                           _jvz, _jax, _jay, _jaz, _jjx, _jjy, _jjz};
    float* oh[6] = {_isx, _isy, _isz, _icx, _icy, _icz};
    tupan_host_pairwise_f32(TUPAN_K_SNAP_CRACKLE, ni, ih, nj, jh, oh, 0.0f);
}

void tstep_kernel(const unsigned int ni, const float* _im, const float* _irx, const float* _iry,
                  const float* _irz, const float* _ie2, const float* _ivx, const float* _ivy,
                  const float* _ivz, const unsigned int nj, const float* _jm, const float* _jrx,
                  const float* _jry, const float* _jrz, const float* _je2, const float* _jvx,
                  const float* _jvy, const float* _jvz, const float eta, float* _idt_a,
                  float* _idt_b)
{
    const float* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const float* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    float* oh[2] = {_idt_a, _idt_b};
    tupan_host_pairwise_f32(TUPAN_K_TSTEP, ni, ih, nj, jh, oh, eta);
}

void nreg_Xkernel(const unsigned int ni, const float* _im, const float* _irx, const float* _iry,
                  const float* _irz, const float* _ie2, const float* _ivx, const float* _ivy,
                  const float* _ivz, const unsigned int nj, const float* _jm, const float* _jrx,
                  const float* _jry, const float* _jrz, const float* _je2, const float* _jvx,
                  const float* _jvy, const float* _jvz, const float dt, float* _idrx,
                  float* _idry, float* _idrz, float* _iax, float* _iay, float* _iaz, float* _iu)
{
    const float* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const float* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    float* oh[7] = {_idrx, _idry, _idrz, _iax, _iay, _iaz, _iu};
    tupan_host_pairwise_f32(TUPAN_K_NREG_X, ni, ih, nj, jh, oh, dt);
}

void nreg_Vkernel(const unsigned int ni, const float* _im, const float* _ivx, const float* _ivy,
                  const float* _ivz, const float* _iax, const float* _iay, const float* _iaz,
                  const unsigned int nj, const float* _jm, const float* _jvx, const float* _jvy,
                  const float* _jvz, const float* _jax, const float* _jay, const float* _jaz,
                  const float dt, float* _idvx, float* _idvy, float* _idvz, float* _ik)
{
    const float* ih[7] = {_im, _ivx, _ivy, _ivz, _iax, _iay, _iaz};
    const float* jh[7] = {_jm, _jvx, _jvy, _jvz, _jax, _jay, _jaz};
    float* oh[4] = {_idvx, _idvy, _idvz, _ik};
    tupan_host_pairwise_f32(TUPAN_K_NREG_V, ni, ih, nj, jh, oh, dt);
}

void pnacc_kernel(const unsigned int ni, const float* _im, const float* _irx, const float* _iry,
                  const float* _irz, const float* _ie2, const float* _ivx, const float* _ivy,
                  const float* _ivz, const unsigned int nj, const float* _jm, const float* _jrx,
                  const float* _jry, const float* _jrz, const float* _je2, const float* _jvx,
                  const float* _jvy, const float* _jvz, const unsigned int order, const float inv1,
                  const float inv2, const float inv3, const float inv4, const float inv5,
                  const float inv6, const float inv7, float* _ipnax, float* _ipnay, float* _ipnaz)
{
    const float* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const float* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    float* oh[3] = {_ipnax, _ipnay, _ipnaz};
    const float inv[7] = {inv1, inv2, inv3, inv4, inv5, inv6, inv7};
    tupan_host_pnacc_f32(ni, ih, nj, jh, order, inv, oh);
}

void sakura_kernel(const unsigned int ni, const float* _im, const float* _irx,
                   const float* _iry, const float* _irz, const float* _ie2, const float* _ivx,
                   const float* _ivy, const float* _ivz, const unsigned int nj, const float* _jm,
                   const float* _jrx, const float* _jry, const float* _jrz, const float* _je2,
                   const float* _jvx, const float* _jvy, const float* _jvz, const float dt,
                   const int flag, float* _idrx, float* _idry, float* _idrz, float* _idvx,
                   float* _idvy, float* _idvz)
{
    const float* ih[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    const float* jh[8] = {_jm, _jrx, _jry, _jrz, _je2, _jvx, _jvy, _jvz};
    float* oh[6] = {_idrx, _idry, _idrz, _idvx, _idvy, _idvz};
    tupan_host_sakura_f32(ni, ih, nj, jh, dt, flag, oh);
}

void kepler_solver_kernel(const float* _im, const float* _irx, const float* _iry,
                          const float* _irz, const float* _ie2, const float* _ivx,
                          const float* _ivy, const float* _ivz, const float dt, float* _ir1x,
                          float* _ir1y, float* _ir1z, float* _iv1x, float* _iv1y, float* _iv1z)
{
    const float* in[8] = {_im, _irx, _iry, _irz, _ie2, _ivx, _ivy, _ivz};
    float* out[6] = {_ir1x, _ir1y, _ir1z, _iv1x, _iv1y, _iv1z};
    tupan_host_kepler_f32(in, dt, out);
}

const char* tupan_f32_last_error(void) { return tupan_last_error(); }

}  // extern "C"
