// vk_noalloc.hip -- vk_fft_c2r_3d_device_noalloc: the 3D C2R of
// vk_fft_c2r_3d_device with no plan and no device allocation.
//
// Its intermediates live in `scratch`, a file-scope device array of fixed
// size that every GPU the library runs on holds until the process ends:
//   [0, TW_BYTES)                 twiddle tables of every supported length
//   [TW_BYTES, +BUF_BYTES)        t1 (nx*ny*hp complex doubles at most)
//   [TW_BYTES+BUF_BYTES, +BUF_BYTES)  t2
// In host code `scratch` names a host-side stand-in, not a device address: the
// device address of each GPU's copy comes from hipGetSymbolAddress, looked up
// once per device (with that device current) at the first call there, and
// only that address is used afterwards.  Every call on a GPU shares that GPU's
// array, so calls that overlap on one GPU race.  No pass reads a scratch
// element it has not written in the same call, so nothing is cleared.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/vonkarman_hip.h"
#include "vk_launch.h"

namespace vk {
int set_error(int code, const char *msg);
int fft_c2r_launch(int nx, int ny, int nz, int hp, const void *spec, void *t1, void *t2, double *real,
                   const void *tw_x, const void *tw_y, const void *tw_z, hipStream_t st);
std::vector<double> host_twiddles(int L);
bool fft_length_ok(int n);
}  // namespace vk

namespace {

constexpr int kLengths[] = {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768};
constexpr size_t TW_BYTES = 64 * 1024;              // sum of 2 L doubles over kLengths: 39,552 B
constexpr size_t BUF_BYTES = 32ull * 1024 * 1024;   // one intermediate: nx*ny*hp*16 B at most
constexpr size_t SCRATCH_BYTES = TW_BYTES + 2 * BUF_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};  // device address of `scratch` on each device

size_t twiddle_offset(int L) {
    size_t off = 0;
    for (int l : kLengths) {
        if (l == L) return off;
        off += 2 * (size_t)l * sizeof(double);
    }
    return (size_t)-1;
}

// Set-up on the current device, once: the array's device address and the
// twiddle tables (computed on the host in extended precision, as the plans'
// are, so both entry points give the same bits).
int device_scratch(unsigned char **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return vk::set_error(VK_ERR_HIP, "vk_fft_c2r_3d_device_noalloc: no current HIP device");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return vk::set_error(VK_ERR_HIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
        auto *base = static_cast<unsigned char *>(p);
        for (int L : kLengths) {
            const std::vector<double> w = vk::host_twiddles(L);
            e = hipMemcpy(base + twiddle_offset(L), w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice);
            if (e != hipSuccess)
                return vk::set_error(VK_ERR_HIP, (std::string("twiddle upload: ") + hipGetErrorString(e)).c_str());
        }
        g_ptr[dev] = base;
    }
    *out = g_ptr[dev];
    return VK_OK;
}

}  // namespace

extern "C" int vk_fft_c2r_3d_device_noalloc(int nx, int ny, int nz, const double *d_in, double *d_out, void *stream) {
    if (!d_in || !d_out) return vk::set_error(VK_ERR_ARG, "vk_fft_c2r_3d_device_noalloc: null argument");
    if (!vk::fft_length_ok(nx) || !vk::fft_length_ok(ny) || !vk::fft_length_ok(nz))
        return vk::set_error(VK_ERR_UNSUPPORTED, "vk_fft_c2r_3d_device_noalloc: lengths must be 2^a 3^b in [8, 768]");
    const int hz = nz / 2 + 1;
    const int row = VK_HP_ALIGN_BYTES / 16;
    const int hp = (hz + row - 1) / row * row;
    const size_t need = (size_t)nx * ny * hp * 16;
    if (need > BUF_BYTES)
        return vk::set_error(VK_ERR_ARG, ("vk_fft_c2r_3d_device_noalloc: grid needs " + std::to_string(need) +
                                          " B per intermediate, above the fixed " + std::to_string(BUF_BYTES) +
                                          " B (nx*ny*hp*16 B; e.g. 128^3 fits)").c_str());
    unsigned char *base = nullptr;
    if (int r = device_scratch(&base)) return r;
    unsigned char *t1 = base + TW_BYTES, *t2 = base + TW_BYTES + BUF_BYTES;
    // the launch sequence reports its own error code and message
    return vk::fft_c2r_launch(nx, ny, nz, hp, d_in, t1, t2, d_out, base + twiddle_offset(nx), base + twiddle_offset(ny),
                              base + twiddle_offset(nz), static_cast<hipStream_t>(stream));
}
