// cmi_ionize.hip -- per-cell ionization-equilibrium update in FP64
// (DensitySubGrid::compute_neutral_fraction, DensitySubGrid.hpp:1628-1658),
// the FP32 opacity refresh, and the L1-change reduction.
//
// HBM-bound streaming kernel: per cell it reads J (8 B), n (8 B), x_n (8 B)
// and writes x_n (8 B), J = 0 (8 B) and, in FP32 mode, kappa (4 B).
#include "cmi_internal.hpp"

namespace cmi {

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
  }
  return s;
}

__global__ void __launch_bounds__(256) ionize_kernel(
    const double *__restrict__ n, double *__restrict__ xn,
    double *__restrict__ J, float *__restrict__ kappa, size_t ncell,
    double jfac, double jscale, double alpha, int form, double kappa_scale,
    int write_kappa, double *__restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ double sh[4];
  double acc = 0.;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncell;
       i += (size_t)gridDim.x * blockDim.x) {
    const double Jc = (jscale == 1.) ? J[i] : J[i] * jscale;
    const double jH = jfac * Jc;
    const double ntot = n[i];
    double x;
    if (jH > 0.) {
      const double aa = 0.5 * jH / (ntot * alpha);
      if (form == 0) {
        const double bb = 2. / aa;
        const double cc = sqrt(bb + 1.);
        x = 1. + aa * (1. - cc);
      } else {
        const double b = 2. * aa;
        x = 2. / (2. + b + sqrt(b * (b + 4.)));
      }
    } else {
      x = 1.;
    }
    acc += fabs(x - xn[i]);
    xn[i] = x;
    J[i] = 0.;
    if (write_kappa) kappa[i] = (float)(kappa_scale * ntot * x);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void kappa_kernel(const double *__restrict__ n,
                             const double *__restrict__ xn,
                             float *__restrict__ kappa, size_t ncell,
                             double kappa_scale) {
#pragma clang fp contract(off)
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncell;
       i += (size_t)gridDim.x * blockDim.x) {
    kappa[i] = (float)(kappa_scale * n[i] * xn[i]);
  }
}

// A store, not a scaling: freshly allocated memory may hold NaN or Inf bit
// patterns, which a multiplication by zero would keep.
__global__ void fill_kernel(double *__restrict__ p, size_t n, double value) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = value;
}

hipError_t launch_ionize(const FieldsDev &f, size_t ncell, double jfac,
                         double jscale, double alpha, int form,
                         double kappa_scale, int write_kappa, double *partial,
                         int blocks, hipStream_t st) {
  hipLaunchKernelGGL(ionize_kernel, dim3(blocks), dim3(256), 0, st, f.n, f.xn,
                     f.J, f.kappa, ncell, jfac, jscale, alpha, form,
                     kappa_scale, write_kappa, partial);
  return hipGetLastError();
}

hipError_t launch_kappa(const FieldsDev &f, size_t ncell, double kappa_scale,
                        hipStream_t st) {
  const unsigned blocks =
      (unsigned)((ncell + 255) / 256 < 4096 ? (ncell + 255) / 256 : 4096);
  hipLaunchKernelGGL(kappa_kernel, dim3(blocks ? blocks : 1), dim3(256), 0, st,
                     f.n, f.xn, f.kappa, ncell, kappa_scale);
  return hipGetLastError();
}

hipError_t launch_fill(double *p, size_t n, double value, hipStream_t st) {
  const unsigned blocks =
      (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(fill_kernel, dim3(blocks ? blocks : 1), dim3(256), 0, st,
                     p, n, value);
  return hipGetLastError();
}

}  // namespace cmi
