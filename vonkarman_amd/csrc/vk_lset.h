// vk_lset.h -- per-length launcher templates.  Their definitions live in
// vk_kernels.inl and are explicitly instantiated once per supported length by
// vk_inst.hip (-DVK_L=<len>); the dispatchers only see these declarations.
#pragma once
#include "vk_launch.h"

namespace vk {
template <int L>
int launch_inv_L(int mode, int lin, Prec ip, Prec wp, const void *in, void *out, const void *tw,
                 const Strided &g, int n, double ks, double scale, hipStream_t st);
template <int L>
int launch_fwd_L(int mode, int lout, int nf, Prec wp, const void *in, void *out, const void *tw,
                 const Strided &g, int n, double ks, double scale, const StageArgs *sa, hipStream_t st);
template <int L>
int launch_fwd_inv_L(int lout, Prec wp, const void *in, void *aout, const void *tw, const Strided &g,
                     const Strided &gi, int n, double ks, const StageArgs *sa, hipStream_t st);
template <int L>
int launch_z_nl_L(Prec wp, const void *in, void *out, long long lines, int hp, long long fs_in,
                  long long fs_out, double s_in, double s_out, const void *tw, hipStream_t st);
template <int L>
int launch_z_vjp_L(Prec wp, const void *in, void *out, long long lines, int hp, long long fs_in,
                   long long fs_out, double s_u, double s_o, const void *tw, hipStream_t st);
template <int L>
int launch_z_max_L(Prec wp, const void *in, long long lines, int hp, long long fs, double s_in,
                   const void *tw, unsigned long long *slots, hipStream_t st);
template <int L>
int launch_z_c2r_L(const void *in, double *out, long long lines, int hp, double scale, const void *tw,
                   hipStream_t st);
template <int L>
int launch_z_r2c_L(const double *in, void *out, long long lines, int hp, const void *tw, hipStream_t st);
}  // namespace vk

// the supported FFT lengths: 2^a 3^b from 8 to 768 (grids N = 8..512 and
// their 3/2-padded sizes M = 12..768)
#define VK_FOR_EACH_LENGTH(X) \
    X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)
