// rp_noalloc.hip -- the Poisson solve of FFT::doFFT (src/fft.cpp:112-177) on
// device buffers, for callers that cannot allocate device memory (code under
// HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// It runs the kernels of the placement context's solve (row DCT, column
// spectral pass, row inverse) on a scratch region that is a file-scope
// device array of fixed size: the FFT work rows, the field, the twiddle and
// frequency tables, the overflow partials and the controller word the
// kernels read.  Each GPU holds its own copy of the array from the first
// call on it until the process ends; its address on the current device is
// looked up once per device with hipGetSymbolAddress.  Every call on a GPU
// shares that GPU's array, so calls that overlap on one GPU (on different
// streams or host threads) race; calls on one stream are ordered.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "replace/rp_capi.h"
#include "rp_dev_struct.h"

hipError_t rp_launch_poisson(const RpDev& d, hipStream_t s, int want_phi, int from_rho, const RpMfmaDev* mf);

namespace {

constexpr int kMaxSide = RP_MAX_FFT;  // 2048 bins per side
constexpr size_t kAlign = 256;
constexpr size_t up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Scratch layout for a grid of at most RP_NOALLOC_MAX_BINS bins.
constexpr size_t kB = RP_NOALLOC_MAX_BINS;
constexpr size_t kOffT = 0;                                     // float[B]: row DCT
constexpr size_t kOffY1 = kOffT + up(sizeof(float) * kB);       // float[B]: column pass, phi / Ex part
constexpr size_t kOffY2 = kOffY1 + up(sizeof(float) * kB);      // float[B]: column pass, Ey part
constexpr size_t kOffField = kOffY2 + up(sizeof(float) * kB);   // float2[B]: (Ex, Ey)
constexpr size_t kOffTwx = kOffField + up(sizeof(float2) * kB); // float2[max/2]
constexpr size_t kOffTw2x = kOffTwx + up(sizeof(float2) * kMaxSide / 2);  // float2[max]
constexpr size_t kOffTwy = kOffTw2x + up(sizeof(float2) * kMaxSide);
constexpr size_t kOffTw2y = kOffTwy + up(sizeof(float2) * kMaxSide / 2);
constexpr size_t kOffWx = kOffTw2y + up(sizeof(float2) * kMaxSide);       // float[max] x4
constexpr size_t kOffWx2 = kOffWx + up(sizeof(float) * kMaxSide);
constexpr size_t kOffWy = kOffWx2 + up(sizeof(float) * kMaxSide);
constexpr size_t kOffWy2 = kOffWy + up(sizeof(float) * kMaxSide);
constexpr size_t kOffPart = kOffWy2 + up(sizeof(float) * kMaxSide);       // double[max/2]
constexpr size_t kOffState = kOffPart + up(sizeof(double) * kMaxSide / 2);
constexpr size_t kBytes = kOffState + up(sizeof(RpState));

}  // namespace

// The scratch array (one per GPU).  Host code never uses this name as an
// address: it is a host-side stand-in; the device address comes from
// hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[kBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];  // scratch address on each device, looked up once

// The address of this device's scratch array (nullptr on failure).
void* scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

// Twiddle and frequency tables of the grid (the placement context builds
// the same tables on the host, FFT::init src/fft.cpp:78-90), and the
// controller word the kernels check, set to "running".
__global__ void k_np_tables(RpDev d) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const double pi = 3.141592653589793238462;
  float2* twx = const_cast<float2*>(d.twx);
  float2* tw2x = const_cast<float2*>(d.tw2x);
  float2* twy = const_cast<float2*>(d.twy);
  float2* tw2y = const_cast<float2*>(d.tw2y);
  float* wx = const_cast<float*>(d.wx);
  float* wx2 = const_cast<float*>(d.wx2);
  float* wy = const_cast<float*>(d.wy);
  float* wy2 = const_cast<float*>(d.wy2);
  if (k < d.nbx / 2) {
    const double a = -2.0 * pi * (double)k / (double)d.nbx;
    twx[k] = make_float2((float)cos(a), (float)sin(a));
  }
  if (k < d.nby / 2) {
    const double a = -2.0 * pi * (double)k / (double)d.nby;
    twy[k] = make_float2((float)cos(a), (float)sin(a));
  }
  if (k < d.nbx) {
    const double a = pi * (double)k / (2.0 * (double)d.nbx);
    tw2x[k] = make_float2((float)cos(a), (float)-sin(a));
    const float w = (float)(pi * (double)(float)k / (double)(float)d.nbx);
    wx[k] = w;
    wx2[k] = w * w;
  }
  if (k < d.nby) {
    const double a = pi * (double)k / (2.0 * (double)d.nby);
    tw2y[k] = make_float2((float)cos(a), (float)-sin(a));
    const float w = (float)(pi * (double)(float)k / (double)(float)d.nby * (double)(float)d.bsy /
                            (double)(float)d.bsx);
    wy[k] = w;
    wy2[k] = w * w;
  }
  if (k == 0) d.st->phase = 0;
}

// (Ex, Ey) pairs into the caller's two arrays.
__global__ void k_np_split(const float2* __restrict__ field, float* __restrict__ ex, float* __restrict__ ey,
                           int n) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const float2 f = field[b];
  ex[b] = f.x;
  ey[b] = f.y;
}

bool pow2_side(int v) { return v >= 4 && v <= kMaxSide && (v & (v - 1)) == 0; }

}  // namespace

extern "C" int rp_poisson_solve_noalloc(const float* density, int32_t nx, int32_t ny, int32_t bin_size_x,
                                        int32_t bin_size_y, float* phi, float* ex, float* ey, void* stream) {
  if (!density || !ex || !ey || bin_size_x <= 0 || bin_size_y <= 0) return RP_ERR_ARG;
  if (!pow2_side(nx) || !pow2_side(ny) || (int64_t)nx * ny > (int64_t)RP_NOALLOC_MAX_BINS) return RP_ERR_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return RP_ERR_NO_DEVICE;
  unsigned char* base = static_cast<unsigned char*>(scratch_ptr(dev));
  if (!base) return RP_ERR_HIP;
  hipStream_t s = static_cast<hipStream_t>(stream);
  RpDev d;
  std::memset(&d, 0, sizeof d);
  d.nbx = nx;
  d.nby = ny;
  d.bsx = bin_size_x;
  d.bsy = bin_size_y;
  d.rho = const_cast<float*>(density);  // read only (the solve starts from the given density)
  d.phi = phi;
  d.t = reinterpret_cast<float*>(base + kOffT);
  d.y1 = reinterpret_cast<float*>(base + kOffY1);
  d.y2 = reinterpret_cast<float*>(base + kOffY2);
  d.field = reinterpret_cast<float2*>(base + kOffField);
  d.twx = reinterpret_cast<const float2*>(base + kOffTwx);
  d.tw2x = reinterpret_cast<const float2*>(base + kOffTw2x);
  d.twy = reinterpret_cast<const float2*>(base + kOffTwy);
  d.tw2y = reinterpret_cast<const float2*>(base + kOffTw2y);
  d.wx = reinterpret_cast<const float*>(base + kOffWx);
  d.wx2 = reinterpret_cast<const float*>(base + kOffWx2);
  d.wy = reinterpret_cast<const float*>(base + kOffWy);
  d.wy2 = reinterpret_cast<const float*>(base + kOffWy2);
  d.part_bin = reinterpret_cast<double*>(base + kOffPart);
  d.st = reinterpret_cast<RpState*>(base + kOffState);
  const int side = nx > ny ? nx : ny;
  hipLaunchKernelGGL(k_np_tables, dim3((side + 255) / 256), dim3(256), 0, s, d);
  if (hipGetLastError() != hipSuccess) return RP_ERR_HIP;
  if (rp_launch_poisson(d, s, phi ? 1 : 0, 1, nullptr) != hipSuccess) return RP_ERR_HIP;
  const int nb = nx * ny;
  hipLaunchKernelGGL(k_np_split, dim3((nb + 255) / 256), dim3(256), 0, s, (const float2*)d.field, ex, ey, nb);
  return hipGetLastError() == hipSuccess ? RP_OK : RP_ERR_HIP;
}
