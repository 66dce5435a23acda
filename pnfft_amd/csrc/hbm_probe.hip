// Achievable HBM bandwidth on the device, for grading the HBM-bound kernels of
// the library (FFT passes, slot fold): hand-written stream kernels over two
// 1 GiB buffers, timed with HIP events.  Built as its own small library
// (libpnfft_probe.so); nothing in libpnfft_amd.so uses it.
#include <hip/hip_runtime.h>

namespace {

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr long kBytes = 1l << 30;
constexpr int kEdge = 512;  // the pattern kernels see the buffers as [512][512][512] complex floats

// one 16-byte element per thread; NT: nontemporal loads and stores
template <bool NT>
__global__ __launch_bounds__(256) void copy_k(const f32x4* __restrict__ a, f32x4* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  if constexpr (NT)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
  else
    b[i] = a[i];
}

// eight 16-byte elements per thread, summed; the sum is stored only if it has
// a value the buffer (all ones) cannot give, so the loads are kept
__global__ __launch_bounds__(256) void read_k(const f32x4* __restrict__ a, float* out, long n) {
  const long i0 = (long)blockIdx.x * 256 * 8 + threadIdx.x;
  float s = 0.f;
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const long i = i0 + 256l * u;
    if (i < n) {
      const f32x4 v = a[i];
      s += v.x + v.y + v.z + v.w;
    }
  }
  if (s == -1.f) out[0] = s;
}

__global__ __launch_bounds__(256) void write_k(f32x4* __restrict__ b, long n) {
  const long i0 = (long)blockIdx.x * 256 * 4 + threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const long i = i0 + 256l * u;
    if (i < n) b[i] = f32x4{2.f, 2.f, 2.f, 2.f};
  }
}

// the axis-0 column pass's access pattern: pencils along axis 0 (elements
// 512 * 512 apart), 32 consecutive pencils per 512-thread block, one 8-byte
// element per lane and 32 elements per thread, all loaded before the stores
__global__ __launch_bounds__(512) void cols_k(const f32x2* __restrict__ a, f32x2* __restrict__ b) {
  const int c = threadIdx.x % 32, j0 = threadIdx.x / 32;
  const long base = (long)blockIdx.y * kEdge + (long)blockIdx.x * 32 + c;
  const long stride = (long)kEdge * kEdge;
  f32x2 v[32];
#pragma unroll
  for (int i = 0; i < 32; i++) v[i] = a[base + (long)(j0 + 16 * i) * stride];
#pragma unroll
  for (int i = 0; i < 32; i++) b[base + (long)(j0 + 16 * i) * stride] = v[i];
}

// the 512-point rows pass's access pattern: one row per wave, element lane + 64 q
__global__ __launch_bounds__(256) void rows_k(const f32x2* __restrict__ a, f32x2* __restrict__ b) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long row = (long)blockIdx.x * 4 + wave;
  f32x2 v[8];
#pragma unroll
  for (int q = 0; q < 8; q++) v[q] = a[row * kEdge + lane + 64 * q];
#pragma unroll
  for (int q = 0; q < 8; q++) b[row * kEdge + lane + 64 * q] = v[q];
}

}  // namespace

// GB/s of: [0] copy, [1] copy with nontemporal accesses, [2] read, [3] write,
// [4] copy with the axis-0 column pattern, [5] copy with the rows pattern
// (bytes read plus bytes written per call, over the mean of `reps` timed calls
// after one warm-up call each).  Returns 0, or -1 when a HIP call fails.
extern "C" int pnfft_amd_hbm_probe(double* out, int reps) {
  static_assert((long)kEdge * kEdge * kEdge * 8 == kBytes, "pattern kernels cover the buffers");
  f32x4 *a = nullptr, *b = nullptr;
  float* o = nullptr;
  const long n = kBytes / 16;
  if (hipMalloc(&a, kBytes) != hipSuccess || hipMalloc(&b, kBytes) != hipSuccess || hipMalloc(&o, 64) != hipSuccess) {
    (void)hipFree(a);
    (void)hipFree(b);
    return -1;
  }
  int rc = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMemset(a, 0, kBytes) != hipSuccess || hipMemset(b, 0, kBytes) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = -1;
  auto timed = [&](double traffic, auto launch) -> double {
    launch();
    if (hipDeviceSynchronize() != hipSuccess) return -1.0;
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++) launch();
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) return -1.0;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return traffic / (ms / reps * 1e-3) / 1e9;
  };
  if (rc == 0) {
    const unsigned g1 = (unsigned)((n + 255) / 256);
    const f32x2* a2 = reinterpret_cast<const f32x2*>(a);
    f32x2* b2 = reinterpret_cast<f32x2*>(b);
    out[0] = timed(2.0 * kBytes, [&] { hipLaunchKernelGGL(copy_k<false>, dim3(g1), dim3(256), 0, 0, a, b, n); });
    out[1] = timed(2.0 * kBytes, [&] { hipLaunchKernelGGL(copy_k<true>, dim3(g1), dim3(256), 0, 0, a, b, n); });
    out[2] = timed(1.0 * kBytes,
                   [&] { hipLaunchKernelGGL(read_k, dim3((unsigned)((n + 2047) / 2048)), dim3(256), 0, 0, a, o, n); });
    out[3] = timed(1.0 * kBytes,
                   [&] { hipLaunchKernelGGL(write_k, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, 0, b, n); });
    out[4] = timed(2.0 * kBytes,
                   [&] { hipLaunchKernelGGL(cols_k, dim3(kEdge / 32, kEdge), dim3(512), 0, 0, a2, b2); });
    out[5] = timed(2.0 * kBytes,
                   [&] { hipLaunchKernelGGL(rows_k, dim3(kEdge * kEdge / 4), dim3(256), 0, 0, a2, b2); });
    for (int k = 0; k < 6; k++)
      if (out[k] < 0) rc = -1;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(o);
  return rc;
}
