// vk_inst.hip -- explicit instantiation of every kernel for one FFT length.
// Built once per length with -DVK_L=<len> so lengths compile in parallel.
#include "vk_kernels.inl"

#ifndef VK_L
#error "compile with -DVK_L=<length>"
#endif

namespace vk {
template int launch_inv_L<VK_L>(int, int, Prec, Prec, const void *, void *, const void *, const Strided &, int,
                                double, double, hipStream_t);
template int launch_fwd_L<VK_L>(int, int, int, Prec, const void *, void *, const void *, const Strided &, int,
                                double, double, const StageArgs *, hipStream_t);
template int launch_fwd_inv_L<VK_L>(int, Prec, const void *, void *, const void *, const Strided &, const Strided &,
                                    int, double, const StageArgs *, hipStream_t);
template int launch_z_nl_L<VK_L>(Prec, const void *, void *, long long, int, long long, long long, double, double,
                                 const void *, hipStream_t);
template int launch_z_vjp_L<VK_L>(Prec, const void *, void *, long long, int, long long, long long, double, double,
                                  const void *, hipStream_t);
template int launch_z_max_L<VK_L>(Prec, const void *, long long, int, long long, double, const void *,
                                  unsigned long long *, hipStream_t);
template int launch_z_c2r_L<VK_L>(const void *, double *, long long, int, double, const void *, hipStream_t);
template int launch_z_r2c_L<VK_L>(const double *, void *, long long, int, const void *, hipStream_t);
}  // namespace vk
