// batch.hip -- many 2-D fan-beam slices of uneven size, view count and basis
// (B0 / B1 / SF) in one persistent launch per direction.
//
// The reference projects one slice at a time with one launch per view
// (layers/A_B1.cu:512-541).  Here a slice batch is a work queue:
//   A    items are waves of 64 bins of one view of one slice (the unit of
//        fwd_bin_wave / sf_fwd_wave); each wave of the persistent grid pulls
//        the next item from a device counter until the queue is empty.
//   A^T  items are 16x16 pixel tiles of one slice (bwd_pix_tile /
//        sf_bwd_tile); each 256-thread block pulls the next tile.
// Items are ordered slice by slice, slices by decreasing per-item cost (the
// longest-processing-time-first rule), so the big slices start first and
// the small ones fill the tail.  Every item runs the same device function as
// the per-slice kernels, so results equal their unsplit launches bit for bit
// (a split launch's FP64 partial sums can differ by one FP32 rounding).  Counters are
// cleared by a one-block kernel launched just before (no memset node, so the
// pair can be captured in a graph and replayed).  The queue ends when the
// counter passes the item count, which every wave / block reaches.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "proj2d_dev.hpp"
#include "sf2d_dev.hpp"

namespace sgdct {

__device__ __forceinline__ int find_slice(const long long* prefix, int n, long long it) {
  // largest k with prefix[k] <= it (prefix[0] = 0, prefix[n] = total)
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= it) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void batch_reset_kernel(unsigned int* counters) {
  if (threadIdx.x < 2) counters[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void batch_fwd_kernel(BatchPlan P, const float* __restrict__ img,
                                                        float* __restrict__ sino) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    unsigned int mine = 0;
    if (lane == 0) mine = atomicAdd(&P.counters[0], 1u);
    // lane 0's value, wave-uniform for the compiler: the slice lookup and
    // descriptor then use scalar loads and registers (through a lane shuffle
    // the kernel held 100 VGPRs, with this 92)
    const long long it = (long long)(unsigned int)__builtin_amdgcn_readfirstlane((int)mine);
    if (it >= P.nfwd) break;
    const int k = find_slice(P.fwd_prefix, P.nslice, it);
    const SliceDesc d = P.slices[P.fwd_order[k]];
    const int wave = (int)(it - P.fwd_prefix[k]);
    const float* x = img + d.img_off;
    float* y = sino + d.sino_off;
    if (d.basis == BASIS_B1) fwd_bin_wave<BASIS_B1>(d.g, x, y, wave);
    else if (d.basis == BASIS_B0) fwd_bin_wave<BASIS_B0>(d.g, x, y, wave);
    else sf_fwd_wave(d.g, x, y, wave);
  }
}

__global__ __launch_bounds__(256, 6) void batch_bwd_kernel(BatchPlan P, const float* __restrict__ sino,
                                                        float* __restrict__ img) {
  __shared__ B1Bin recs[kViewChunk][kMaxWindow];  // also holds B0Bin records
  __shared__ int win[2 * kViewChunk];
  __shared__ unsigned int s_it;
  static_assert(sizeof(B1Bin) >= sizeof(B0Bin), "record buffer sized for B1");
  for (;;) {
    if (threadIdx.x == 0) s_it = atomicAdd(&P.counters[1], 1u);
    __syncthreads();
    // the same value in every lane; wave-uniform for the compiler (scalar
    // slice lookup and descriptor)
    const long long it = (long long)(unsigned int)__builtin_amdgcn_readfirstlane((int)s_it);
    __syncthreads();
    if (it >= P.nbwd) break;
    const int k = find_slice(P.bwd_prefix, P.nslice, it);
    const SliceDesc d = P.slices[P.bwd_order[k]];
    const int tile = (int)(it - P.bwd_prefix[k]);
    const int tiles_x = (d.g.nx + kTile - 1) / kTile;
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const float* y = sino + d.sino_off;
    float* x = img + d.img_off;
    if (d.basis == BASIS_B1) {
      if (d.lds) bwd_pix_tile<BASIS_B1, true>(d.g, y, x, tx, ty, recs, win);
      else bwd_pix_tile<BASIS_B1, false>(d.g, y, x, tx, ty, recs, win);
    } else if (d.basis == BASIS_B0) {
      B0Bin(*r0)[kMaxWindow] = reinterpret_cast<B0Bin(*)[kMaxWindow]>(&recs[0][0]);
      if (d.lds) bwd_pix_tile<BASIS_B0, true>(d.g, y, x, tx, ty, r0, win);
      else bwd_pix_tile<BASIS_B0, false>(d.g, y, x, tx, ty, r0, win);
    } else {
      sf_bwd_tile(d.g, y, x, tx, ty);
    }
    __syncthreads();
  }
}

hipError_t launch_batch_forward(const BatchPlan& P, int blocks, const float* img, float* sino, hipStream_t s) {
  batch_reset_kernel<<<1, 64, 0, s>>>(P.counters);
  batch_fwd_kernel<<<blocks, 256, 0, s>>>(P, img, sino);
  return hipGetLastError();
}

hipError_t launch_batch_backward(const BatchPlan& P, int blocks, const float* sino, float* img, hipStream_t s) {
  batch_reset_kernel<<<1, 64, 0, s>>>(P.counters);
  batch_bwd_kernel<<<blocks, 256, 0, s>>>(P, sino, img);
  return hipGetLastError();
}

}  // namespace sgdct
