// HIP STREAM for gfx950: the measured-peak denominator of the roofline, as the reference measures
// its peak with bench/stream/stream.cu (FP64 arrays moved as 16 B vectors, copy / scale / add /
// triad, the best of several repetitions). One 16 B element per thread.
#pragma once

#include <hip/hip_runtime.h>

namespace mxe {

__global__ __launch_bounds__(256) void k_stream_init(double2* a, double2* b, double2* c, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  a[i] = make_double2(1.0, 1.0);
  b[i] = make_double2(2.0, 2.0);
  c[i] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_stream_copy(const double2* __restrict__ a, double2* __restrict__ c,
                                                     long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) c[i] = a[i];
}

__global__ __launch_bounds__(256) void k_stream_scale(const double2* __restrict__ c, double2* __restrict__ b,
                                                      double s, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) {
    const double2 v = c[i];
    b[i] = make_double2(s * v.x, s * v.y);
  }
}

__global__ __launch_bounds__(256) void k_stream_add(const double2* __restrict__ a, const double2* __restrict__ b,
                                                    double2* __restrict__ c, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) {
    const double2 x = a[i], y = b[i];
    c[i] = make_double2(x.x + y.x, x.y + y.y);
  }
}

__global__ __launch_bounds__(256) void k_stream_triad(const double2* __restrict__ b, const double2* __restrict__ c,
                                                      double2* __restrict__ a, double s, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) {
    const double2 x = b[i], y = c[i];
    a[i] = make_double2(x.x + s * y.x, x.y + s * y.y);
  }
}

}  // namespace mxe
