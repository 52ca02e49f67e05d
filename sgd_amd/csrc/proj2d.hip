// proj2d.hip -- gfx950 kernels of the 2-D fan-beam footprint projectors.
//
//   fwd_bin_kernel   A   (image -> sinogram).  One lane per (view, bin): the
//                    lane walks the pixels of its bin's strip line by line and
//                    gathers sum_pixel conv * x.  No atomics, deterministic.
//   bwd_pix_kernel   A^T (sinogram -> image).  One lane per pixel, 16x16
//                    pixel tiles per workgroup; per-bin constants and the
//                    sinogram values of the bins the tile sees are staged in
//                    LDS, kViewChunk views at a time.  Deterministic gather.
//   sum_parts_kernel adds the FP64 partial results of split launches in a
//                    fixed order: A may split every bin's walk into line
//                    segments and A^T the views into groups (grid z), so that
//                    a single image fills the GPU; each segment / group
//                    writes its partial sums, and the split is chosen per
//                    projector from one image's launch size.
//   count_kernel     number of footprint evaluations under the reference's
//                    corner rule (the value layers/A_B1.cu:50,643 returns).
//
// Footprint values follow layers/A_B1.cu:62-208 (B1), layers/A_B0.cu:123-229
// (B0), layers/A_SF.cu:48-225 (SF); see footprint.hpp for the form used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "proj2d_dev.hpp"

namespace sgdct {

// grid (waves, batch, nseg), one wave per workgroup: waves of a view differ
// in length, and single-wave workgroups let the dispatcher refill each wave
// slot as soon as it frees (four-wave workgroups waited for their slowest
// wave: B1 A 1.058 -> 0.959 ms at 512 x 512, 720 views, 768 bins).  With
// nseg > 1 every (segment, image) writes FP64 partial sums to
// part[(seg * batch + image) * np * nu + ...], which sum_parts_kernel adds in
// segment order.
template <int BASIS>
__global__ __launch_bounds__(64) void fwd_bin_kernel(Proj2DArgs g, const float* __restrict__ img,
                                                     float* __restrict__ sino, double* __restrict__ part) {
  const int b = blockIdx.y, seg = blockIdx.z, nseg = gridDim.z;
  const size_t nsino = (size_t)g.np * g.nu;
  fwd_bin_wave<BASIS>(g, img + (size_t)b * g.nx * g.ny, sino + b * nsino, blockIdx.x, seg, nseg,
                      nseg > 1 ? part + ((size_t)seg * gridDim.y + b) * nsino : nullptr);
}

// out[i] = sum_k part[k n + i], k ascending (deterministic).
__global__ __launch_bounds__(256) void sum_parts_kernel(const double* __restrict__ part, float* __restrict__ out,
                                                        size_t n, int nparts) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < nparts; ++k) acc += part[(size_t)k * n + i];
    out[i] = (float)acc;
  }
}


// Widest bin window over every tile and view (host-side, on the host copy of
// the view table): the LDS-staged A^T kernel is used when it is <= kMaxWindow.
int max_tile_window(const Proj2DArgs& g, const ViewGeo* views_host) {
  int w = 0;
  for (int v = 0; v < g.np; ++v)
    for (int iy0 = 0; iy0 < g.ny; iy0 += kTile)
      for (int ix0 = 0; ix0 < g.nx; ix0 += kTile) {
        int lo, n;
        tile_window(g, views_host[v], ix0, iy0, &lo, &n, 1 << 30);
        w = n > w ? n : w;
      }
  return w;
}

// grid (tiles x, tiles y, batch * ngroup): view group k = blockIdx.z % ngroup
// covers views [k vpg, (k + 1) vpg); with ngroup > 1 it writes FP64 partial
// images to part[(k * batch + image) * nx * ny + ...] for sum_parts_kernel.
template <int BASIS, bool USE_LDS>
__global__ __launch_bounds__(256, 8) void bwd_pix_kernel(Proj2DArgs g, const float* __restrict__ sino,
                                                      float* __restrict__ img, int ngroup, int vpg,
                                                      double* __restrict__ part) {
  using Bin = typename Traits<BASIS>::Bin;
  __shared__ Bin lds[USE_LDS ? kViewChunk : 1][kMaxWindow];
  __shared__ int win[2 * kViewChunk];
  const int k = blockIdx.z % ngroup, b = blockIdx.z / ngroup, batch = gridDim.z / ngroup;
  const size_t nimg = (size_t)g.nx * g.ny;
  const int v0 = k * vpg, v1 = min(g.np, v0 + vpg);
  bwd_pix_tile<BASIS, USE_LDS>(g, sino + (size_t)b * g.np * g.nu, img + b * nimg, blockIdx.x, blockIdx.y, lds, win,
                               v0, v1, ngroup > 1 ? part + ((size_t)k * batch + b) * nimg : nullptr);
}

// ---------------------------------------------------------------------------
// Evaluation count under the reference's clamped corner rule.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void count_kernel(Proj2DArgs g, unsigned long long* out) {
  __shared__ unsigned long long part[256];
  const int ix = blockIdx.x * 16 + (threadIdx.x & 15), iy = blockIdx.y * 16 + (threadIdx.x >> 4);
  unsigned long long c = 0;
  if (ix < g.nx && iy < g.ny) {
    for (int v = 0; v < g.np; ++v) {
      const ViewGeo& vg = g.views[v];
      double umin = 1e300, umax = -1e300;
      for (int k = 0; k < 4; ++k) {
        double fx = ix + ((k & 1) ? g.h : -g.h), fy = iy + ((k & 2) ? g.h : -g.h);
        double wx = fx * g.dx + vg.w0x, wy = fy * g.dy + vg.w0y;
        double uu = -(vg.e0x * wy - vg.e0y * wx) / (vg.pux * wy - vg.puy * wx);
        umin = fmin(umin, uu);
        umax = fmax(umax, uu);
      }
      int lo = (int)fmin(fmax(floor(umin), 0.0), (double)(g.nu - 1));
      int hi = (int)fmin(fmax(floor(umax), 0.0), (double)(g.nu - 1));
      c += (unsigned long long)(hi - lo + 1);
    }
  }
  part[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_sum_parts(const double* part, float* out, size_t n, int nparts, hipStream_t s) {
  sum_parts_kernel<<<(unsigned)std::min<size_t>((n + 255) / 256, 4096), 256, 0, s>>>(part, out, n, nparts);
  return hipGetLastError();
}

// nseg > 1 needs part with nseg * batch * np * nu doubles.
hipError_t launch_forward(int basis, const Proj2DArgs& g, const float* img, float* sino, int batch, int nseg,
                          double* part, hipStream_t s) {
  const int waves = g.np * g.nchunk;
  if (nseg < 1 || (nseg > 1 && !part)) return hipErrorInvalidValue;
  dim3 grid(waves, batch, nseg), block(64);
  if (basis == BASIS_B1) fwd_bin_kernel<BASIS_B1><<<grid, block, 0, s>>>(g, img, sino, part);
  else if (basis == BASIS_B0) fwd_bin_kernel<BASIS_B0><<<grid, block, 0, s>>>(g, img, sino, part);
  else return hipErrorInvalidValue;
  if (nseg > 1) {
    const size_t n = (size_t)batch * g.np * g.nu;
    sum_parts_kernel<<<(unsigned)std::min<size_t>((n + 255) / 256, 4096), 256, 0, s>>>(part, sino, n, nseg);
  }
  return hipGetLastError();
}

// ngroup > 1 needs part with ngroup * batch * nx * ny doubles.
hipError_t launch_backward(int basis, const Proj2DArgs& g, bool use_lds, const float* sino, float* img, int batch,
                           int ngroup, double* part, hipStream_t s) {
  if (ngroup < 1 || (ngroup > 1 && !part)) return hipErrorInvalidValue;
  // views per group, whole view chunks
  const int vpg = ((g.np + ngroup - 1) / ngroup + kViewChunk - 1) / kViewChunk * kViewChunk;
  dim3 grid((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile, batch * ngroup), block(256);
  if (basis == BASIS_B1) {
    if (use_lds) bwd_pix_kernel<BASIS_B1, true><<<grid, block, 0, s>>>(g, sino, img, ngroup, vpg, part);
    else bwd_pix_kernel<BASIS_B1, false><<<grid, block, 0, s>>>(g, sino, img, ngroup, vpg, part);
  } else if (basis == BASIS_B0) {
    if (use_lds) bwd_pix_kernel<BASIS_B0, true><<<grid, block, 0, s>>>(g, sino, img, ngroup, vpg, part);
    else bwd_pix_kernel<BASIS_B0, false><<<grid, block, 0, s>>>(g, sino, img, ngroup, vpg, part);
  } else {
    return hipErrorInvalidValue;
  }
  if (ngroup > 1) {
    const size_t n = (size_t)batch * g.nx * g.ny;
    sum_parts_kernel<<<(unsigned)std::min<size_t>((n + 255) / 256, 4096), 256, 0, s>>>(part, img, n, ngroup);
  }
  return hipGetLastError();
}

hipError_t launch_count(const Proj2DArgs& g, unsigned long long* out, hipStream_t s) {
  dim3 grid((g.nx + 15) / 16, (g.ny + 15) / 16), block(256);
  count_kernel<<<grid, block, 0, s>>>(g, out);
  return hipGetLastError();
}

}  // namespace sgdct

