// DispBeamColumn2d/3d over FiberSection2d/3d on gfx950: per-operation
// kernels over a frame group (bodies in xg_frame_impl.h).
#include "xg_frame_impl.h"
#include "xg_kernels.h"

namespace xg {

template <bool D3>
__global__ __launch_bounds__(256) void k_frame_update(FrameDev f, const double* __restrict__ U) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long sec = (long long)blockIdx.x * 4 + wave;
  if (sec >= f.nsec) return;
  section_update<D3>(f, sec, lane, U);
}

// 64 / RW sections of one definition per wavefront (section_rows): rows of
// 16 lanes (four sections, totals by butterfly) or 20 (three sections,
// totals through LDS; held to three waves per SIMD)
template <bool D3>
__global__ __launch_bounds__(256) void k_frame_update_rows(FrameDev f, const double* __restrict__ U) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long w = (long long)blockIdx.x * 4 + wave;
  if (w >= f.nrw) return;
  section_rows<D3, 16>(f, w, lane, U, nullptr);
}
template <bool D3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_frame_update_rows20(
    FrameDev f, const double* __restrict__ U) {
  constexpr int NV = D3 ? 9 : 5;
  __shared__ double red[4 * NV * 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long w = (long long)blockIdx.x * 4 + wave;
  if (w >= f.nrw) return;
  section_rows<D3, 20>(f, w, lane, U, red + wave * NV * 64);
}

template <bool D3>
__global__ __launch_bounds__(256) void k_frame_revert(FrameDev f) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long sec = (long long)blockIdx.x * 4 + wave;
  if (sec >= f.nsec) return;
  section_revert<D3>(f, sec, lane);
}

template <bool D3>
__global__ void k_frame_resid(FrameDev f, double* __restrict__ Re) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= f.nele) return;
  element_resid<D3>(f, e, Re);
}

template <bool D3>
__global__ void k_frame_tangent(FrameDev f, double* __restrict__ Ke) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= f.nele) return;
  element_tangent<D3>(f, e, Ke);
}

__global__ void k_copy_d(long long n, const double* __restrict__ a, double* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ void k_copy_i(long long n, const int* __restrict__ a, int* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

void launch_frame_update(hipStream_t s, const FrameDev& f, bool d3, const double* U, int* err) {
  (void)err;
  if (f.nsec <= 0) return;
  if (f.rsec) {
    const dim3 g(ceil_div(f.nrw, 4)), b(256);
    if (f.rw == 20) {
      if (d3) hipLaunchKernelGGL(k_frame_update_rows20<true>, g, b, 0, s, f, U);
      else hipLaunchKernelGGL(k_frame_update_rows20<false>, g, b, 0, s, f, U);
    } else {
      if (d3) hipLaunchKernelGGL(k_frame_update_rows<true>, g, b, 0, s, f, U);
      else hipLaunchKernelGGL(k_frame_update_rows<false>, g, b, 0, s, f, U);
    }
    return;
  }
  const dim3 g(ceil_div(f.nsec, 4)), b(256);
  if (d3) hipLaunchKernelGGL(k_frame_update<true>, g, b, 0, s, f, U);
  else hipLaunchKernelGGL(k_frame_update<false>, g, b, 0, s, f, U);
}
void launch_frame_revert(hipStream_t s, const FrameDev& f, bool d3) {
  if (f.nsec <= 0) return;
  const dim3 g(ceil_div(f.nsec, 4)), b(256);
  if (d3) hipLaunchKernelGGL(k_frame_revert<true>, g, b, 0, s, f);
  else hipLaunchKernelGGL(k_frame_revert<false>, g, b, 0, s, f);
}
void launch_frame_resid(hipStream_t s, const FrameDev& f, bool d3, double* Re) {
  if (f.nele <= 0) return;
  const dim3 g(ceil_div(f.nele, 128)), b(128);
  if (d3) hipLaunchKernelGGL(k_frame_resid<true>, g, b, 0, s, f, Re);
  else hipLaunchKernelGGL(k_frame_resid<false>, g, b, 0, s, f, Re);
}
void launch_frame_tangent(hipStream_t s, const FrameDev& f, bool d3, double* Ke) {
  if (f.nele <= 0) return;
  const dim3 g(ceil_div(f.nele, 64)), b(64);
  if (d3) hipLaunchKernelGGL(k_frame_tangent<true>, g, b, 0, s, f, Ke);
  else hipLaunchKernelGGL(k_frame_tangent<false>, g, b, 0, s, f, Ke);
}
void launch_copy_d(hipStream_t s, long long n, const double* a, double* b) {
  if (n > 0) hipLaunchKernelGGL(k_copy_d, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, a, b);
}
void launch_copy_i(hipStream_t s, long long n, const int* a, int* b) {
  if (n > 0) hipLaunchKernelGGL(k_copy_i, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, a, b);
}

}  // namespace xg
