// solver.hip -- device kernels of the iterative reconstructions built on A / A^T.
//
// The reference's optimizer/asd_pocs.cu is empty; its notebooks reconstruct
// with pylops CGLS.  These kernels implement the per-pixel parts of SIRT and
// of ASD-POCS (E. Y. Sidky and X. Pan, Phys. Med. Biol. 53 (2008) 4777):
//   * tv_grad_kernel   gradient of the smoothed isotropic total variation
//                      TV(x) = sum_{s,t} sqrt(eps + (x[s,t]-x[s-1,t])^2 + (x[s,t]-x[s,t-1])^2)
//                      (forward differences, zero beyond the image edge);
//   * resid_kernel     r = W (b - A x), W = 1 / (A 1) the SIRT row weights;
//   * update_kernel    x <- max(x + lam * V * c, 0) (positivity optional),
//                      V = 1 / (A^T 1) the SIRT column weights;
//   * axpy_kernel      x <- x + a * g;
//   * sumsq_kernel / tv_step_kernel   one normalised TV-descent step,
//                      x <- x - step * g / ||g|| per image, with ||g|| summed
//                      on the device (no host round trip between steps).
#include <hip/hip_runtime.h>

#include <cmath>

namespace sgdct {

__device__ __forceinline__ float px_(const float* x, int nx, int ny, int i, int j) {
  // image value with the edge replicated (zero difference across the border)
  i = min(max(i, 0), nx - 1);
  j = min(max(j, 0), ny - 1);
  return x[(size_t)j * nx + i];
}

__global__ __launch_bounds__(256) void tv_grad_kernel(const float* __restrict__ x, float* __restrict__ g, int nx,
                                                      int ny, float eps) {
  const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= nx || j >= ny) return;
  x += (size_t)blockIdx.z * nx * ny;
  g += (size_t)blockIdx.z * nx * ny;
  const float c = px_(x, nx, ny, i, j);
  // term centred at (i, j): differences to (i-1, j) and (i, j-1)
  float gx = (i > 0) ? c - px_(x, nx, ny, i - 1, j) : 0.0f;
  float gy = (j > 0) ? c - px_(x, nx, ny, i, j - 1) : 0.0f;
  float out = (gx + gy) * __builtin_amdgcn_rsqf(eps + gx * gx + gy * gy);
  // term centred at (i+1, j): its x-difference contains -x[i,j]
  if (i + 1 < nx) {
    const float r = px_(x, nx, ny, i + 1, j);
    const float dxr = r - c;
    const float dyr = (j > 0) ? r - px_(x, nx, ny, i + 1, j - 1) : 0.0f;
    out -= dxr * __builtin_amdgcn_rsqf(eps + dxr * dxr + dyr * dyr);
  }
  // term centred at (i, j+1): its y-difference contains -x[i,j]
  if (j + 1 < ny) {
    const float u = px_(x, nx, ny, i, j + 1);
    const float dyu = u - c;
    const float dxu = (i > 0) ? u - px_(x, nx, ny, i - 1, j + 1) : 0.0f;
    out -= dyu * __builtin_amdgcn_rsqf(eps + dxu * dxu + dyu * dyu);
  }
  g[(size_t)j * nx + i] = out;
}

__global__ void resid_kernel(const float* __restrict__ b, const float* __restrict__ ax, const float* __restrict__ w,
                             float* __restrict__ r, size_t n, size_t per) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  r[k] = w[k % per] * (b[k] - ax[k]);
}

__global__ void update_kernel(float* __restrict__ x, const float* __restrict__ c, const float* __restrict__ v,
                              float lam, int positivity, size_t n, size_t per) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float y = __builtin_fmaf(lam * v[k % per], c[k], x[k]);
  x[k] = positivity ? fmaxf(y, 0.0f) : y;
}

__global__ void axpy_kernel(float* __restrict__ x, const float* __restrict__ g, float a, size_t n) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  x[k] = __builtin_fmaf(a, g[k], x[k]);
}

// x[k] += a[k / per] * g[k]  (one coefficient per batch entry, on the device)
__global__ void axpy_batch_kernel(float* __restrict__ x, const float* __restrict__ g, const float* __restrict__ a,
                                  size_t n, size_t per) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  x[k] = __builtin_fmaf(a[k / per], g[k], x[k]);
}

// Per-image sums of squares in FP64: block (k, b) adds the elements
// k * 256 + t + j * kTvParts * 256 of image b, then reduces its 256 values
// in a fixed tree; part[b * kTvParts + k] holds the result.
constexpr int kTvParts = 64;

__global__ __launch_bounds__(256) void sumsq_kernel(const float* __restrict__ g, double* __restrict__ part,
                                                    size_t per) {
  __shared__ double red[256];
  const float* gb = g + (size_t)blockIdx.y * per;
  double acc = 0.0;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < per; k += (size_t)kTvParts * 256) {
    const double v = gb[k];
    acc = __builtin_fma(v, v, acc);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * kTvParts + blockIdx.x] = red[0];
}

// x[b] -= step[b] * g[b] / ||g[b]|| (nothing where ||g[b]|| = 0); every block
// of image b adds its kTvParts partial sums in the same order, so all use the
// same norm.  grid (blocks per image, batch).
__global__ __launch_bounds__(256) void tv_step_kernel(float* __restrict__ x, const float* __restrict__ g,
                                                      const double* __restrict__ part,
                                                      const double* __restrict__ step, size_t per) {
  __shared__ float coef;
  const int b = blockIdx.y;
  if (threadIdx.x == 0) {
    double ss = 0.0;
    for (int k = 0; k < kTvParts; ++k) ss += part[(size_t)b * kTvParts + k];
    const double gn = sqrt(ss);
    coef = gn > 0.0 ? (float)(-step[b] / gn) : 0.0f;
  }
  __syncthreads();
  const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= per) return;
  const size_t i = (size_t)b * per + k;
  x[i] = __builtin_fmaf(coef, g[i], x[i]);
}

hipError_t launch_tv_descent(float* x, float* g, double* part, const double* step, int nx, int ny, int batch,
                             float eps, int nsteps, hipStream_t s) {
  const size_t per = (size_t)nx * ny;
  for (int it = 0; it < nsteps; ++it) {
    tv_grad_kernel<<<dim3((nx + 15) / 16, (ny + 15) / 16, batch), dim3(256), 0, s>>>(x, g, nx, ny, eps);
    sumsq_kernel<<<dim3(kTvParts, batch), dim3(256), 0, s>>>(g, part, per);
    tv_step_kernel<<<dim3((unsigned)((per + 255) / 256), batch), dim3(256), 0, s>>>(x, g, part, step, per);
  }
  return hipGetLastError();
}

hipError_t launch_axpy_batch(float* x, const float* g, const float* a, size_t n, size_t per, hipStream_t s) {
  axpy_batch_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(x, g, a, n, per);
  return hipGetLastError();
}

hipError_t launch_tv_grad(const float* x, float* g, int nx, int ny, int batch, float eps, hipStream_t s) {
  tv_grad_kernel<<<dim3((nx + 15) / 16, (ny + 15) / 16, batch), dim3(256), 0, s>>>(x, g, nx, ny, eps);
  return hipGetLastError();
}

hipError_t launch_resid(const float* b, const float* ax, const float* w, float* r, size_t n, size_t per, hipStream_t s) {
  resid_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(b, ax, w, r, n, per);
  return hipGetLastError();
}

hipError_t launch_update(float* x, const float* c, const float* v, float lam, int positivity, size_t n, size_t per,
                         hipStream_t s) {
  update_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(x, c, v, lam, positivity, n, per);
  return hipGetLastError();
}

hipError_t launch_axpy(float* x, const float* g, float a, size_t n, hipStream_t s) {
  axpy_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(x, g, a, n);
  return hipGetLastError();
}

}  // namespace sgdct
