// sf2d.hip -- separable-footprint (SF) fan-beam projector kernels, gfx950.
//
// SF footprint of layers/A_SF.cu:87-225: the four pixel corners (ix +- 1/2,
// iy +- 1/2) are projected to continuous detector coordinates and sorted into
// a trapezoid of unit height; the value for bin u is the trapezoid's area over
// [u, u+1) (gamma, layers/A_SF.cu:48-84) times dx / max(|sin t|, |cos t|),
// t the angle of the source->pixel-centre ray (layers/A_SF.cu:174-177).
//
// Corner projections are formed relative to an integer bin (a nearby edge)
// from the edge functions N_k, so they carry ~1e-7 bin of error instead of
// the ~6e-5 bin that absolute FP32 detector coordinates (up to nu) carry.
#include <hip/hip_runtime.h>

#include <cmath>

#include "sf2d_dev.hpp"

namespace sgdct {

hipError_t launch_sum_parts(const double* part, float* out, size_t n, int nparts, hipStream_t s);

// grid (waves / 4, batch, nseg); partial sums as in fwd_bin_kernel.
__global__ __launch_bounds__(256) void sf_fwd_kernel(Proj2DArgs g, const float* __restrict__ img,
                                                     float* __restrict__ sino, double* __restrict__ part) {
  const int b = blockIdx.y, seg = blockIdx.z, nseg = gridDim.z;
  const size_t nsino = (size_t)g.np * g.nu;
  sf_fwd_wave(g, img + (size_t)b * g.nx * g.ny, sino + b * nsino, blockIdx.x * 4 + (threadIdx.x >> 6), seg, nseg,
              nseg > 1 ? part + ((size_t)seg * gridDim.y + b) * nsino : nullptr);
}

// grid (tiles x, tiles y, batch * ngroup); view groups as in bwd_pix_kernel,
// blocks mapped by xcd_block: SF A^T 0.859 -> 0.840 ms at 512 x 512, 720
// views, 768 bins (B0 / B1's bwd_pix_kernel and Hl's A^T were slower with it).
__global__ __launch_bounds__(256) void sf_bwd_kernel(Proj2DArgs g, const float* __restrict__ sino,
                                                     float* __restrict__ img, int ngroup, int vpg,
                                                     double* __restrict__ part) {
  const uint3 blk = xcd_block();
  const int k = blk.z % ngroup, b = blk.z / ngroup, batch = gridDim.z / ngroup;
  const size_t nimg = (size_t)g.nx * g.ny;
  const int v0 = k * vpg, v1 = min(g.np, v0 + vpg);
  sf_bwd_tile(g, sino + (size_t)b * g.np * g.nu, img + b * nimg, blk.x, blk.y, v0, v1,
              ngroup > 1 ? part + ((size_t)k * batch + b) * nimg : nullptr);
}

hipError_t launch_sf_forward(const Proj2DArgs& g, const float* img, float* sino, int batch, int nseg, double* part,
                             hipStream_t s) {
  if (nseg < 1 || (nseg > 1 && !part)) return hipErrorInvalidValue;
  const int waves = g.np * g.nchunk;
  sf_fwd_kernel<<<dim3((waves + 3) / 4, batch, nseg), dim3(256), 0, s>>>(g, img, sino, part);
  if (nseg > 1) return launch_sum_parts(part, sino, (size_t)batch * g.np * g.nu, nseg, s);
  return hipGetLastError();
}

hipError_t launch_sf_backward(const Proj2DArgs& g, const float* sino, float* img, int batch, int ngroup,
                              double* part, hipStream_t s) {
  if (ngroup < 1 || (ngroup > 1 && !part)) return hipErrorInvalidValue;
  const int vpg = (g.np + ngroup - 1) / ngroup;
  dim3 grid((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile, batch * ngroup);
  sf_bwd_kernel<<<grid, dim3(256), 0, s>>>(g, sino, img, ngroup, vpg, part);
  if (ngroup > 1) return launch_sum_parts(part, img, (size_t)batch * g.nx * g.ny, ngroup, s);
  return hipGetLastError();
}

}  // namespace sgdct

