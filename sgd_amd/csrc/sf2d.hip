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

__global__ __launch_bounds__(256) void sf_fwd_kernel(Proj2DArgs g, const float* __restrict__ img,
                                                     float* __restrict__ sino) {
  sf_fwd_wave(g, img + (size_t)blockIdx.y * g.nx * g.ny, sino + (size_t)blockIdx.y * g.np * g.nu,
              blockIdx.x * 4 + (threadIdx.x >> 6));
}


__global__ __launch_bounds__(256) void sf_bwd_kernel(Proj2DArgs g, const float* __restrict__ sino,
                                                     float* __restrict__ img) {
  sf_bwd_tile(g, sino + (size_t)blockIdx.z * g.np * g.nu, img + (size_t)blockIdx.z * g.nx * g.ny, blockIdx.x,
              blockIdx.y);
}

hipError_t launch_sf_forward(const Proj2DArgs& g, const float* img, float* sino, int batch, hipStream_t s) {
  const int waves = g.np * g.nchunk;
  sf_fwd_kernel<<<dim3((waves + 3) / 4, batch), dim3(256), 0, s>>>(g, img, sino);
  return hipGetLastError();
}

hipError_t launch_sf_backward(const Proj2DArgs& g, const float* sino, float* img, int batch, hipStream_t s) {
  dim3 grid((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile, batch);
  sf_bwd_kernel<<<grid, dim3(256), 0, s>>>(g, sino, img);
  return hipGetLastError();
}

}  // namespace sgdct

