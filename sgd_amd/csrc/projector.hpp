// projector.hpp -- host/device data layout of the 2-D fan-beam projectors.
//
// Geometry convention (utils/geo.cu:232-262 of the reference): view angle b,
// source p = R(b)(-lso + drho, 0), detector centre R(b)(lsd - lso, 0),
// detector step pu = R(b)(0, du); bin u spans detector coordinate [u, u+1),
// i.e. the segment between edge points E_u and E_{u+1},
// E_k = dtv + pu (k - nu/2).  Pixel (ix, iy) has centre
// ((ix - (nx-1)/2) dx, (iy - (ny-1)/2) dy)  (layers/A_B1.cu:96-98).
//
// Edge functions.  N_k(X) = cross(E_k - p, X - p) is zero on the ray through
// edge k and equals |beta(X)| (u(X) - k) for any point X in front of the
// source, u(X) its continuous detector coordinate.  All footprint geometry is
// written with N_k: the reference's corner rule for the bin range of a pixel
// (layers/A_B1.cu:100-151) becomes two affine tests, and the edge-ray offsets
// of layers/A_B1.cu:182-195 become ta = Ka_u N_u, tb = Kb_u N_{u+1}.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sgdct {

enum Basis : int {
  BASIS_B0 = 0, BASIS_B1 = 1, BASIS_SF = 2, BASIS_H0 = 3, BASIS_HL = 4, BASIS_B1_3D = 5, BASIS_SF_3D = 6
};

// Per-view FP64 constants.  With w = X - p = (ix dx + w0x, iy dy + w0y):
//   N_k(ix, iy) = cross(e0 + k pu, w),   beta(ix, iy) = cross(pu, w) (< 0).
struct ViewGeo {
  double e0x, e0y;  // E_0 - p
  double pux, puy;  // detector step
  double w0x, w0y;  // (X - p) at pixel index (0, 0)
  double px, py;    // source
  // FP32 gradients (index units) of N_0 and beta, and the bound
  // db = h (|d beta/d ix| + |d beta/d iy|) used by the per-pixel bin range.
  float gx0, gy0, gx1, gy1, db, pad0, pad1, pad2;
};

// Arguments shared by every 2-D kernel.
struct Proj2DArgs {
  const ViewGeo* views;  // [np]
  const void* bins;      // [np][nu] B0Bin / B1Bin (unused for SF)
  int nx, ny, np, nu;
  double dx, dy;
  float h;               // corner half-extent in index units (1 for B1, 1/2 for B0/SF)
  int nchunk;            // ceil(nu / 64)
};

// Cone-beam B1 (layers/A_B1_3D.cu): the in-plane part reuses Proj2DArgs.
struct Cone3DArgs {
  Proj2DArgs g;     // in-plane geometry and B1 bin records (h = 1)
  int nz, nv, nrc;  // nrc = ceil(nv / 64)
  float vc;         // detector row of the source plane (nv / 2)
  double ldz;       // lsd * dz / dv: v = vc + ldz * (z index offset) / depth
  float cvol;       // dx * dy * dz
  float dz;
  double du;        // |pu|
  float fac;        // separable footprints: lsd^2 dx dy dz / (du dv)
};

// One slice of a batched persistent launch (batch.hip): a 2-D projector's
// arguments and its images' offsets in the packed batch buffers.
struct SliceDesc {
  Proj2DArgs g;
  int basis;               // BASIS_B0, BASIS_B1 or BASIS_SF
  int lds;                 // A^T may stage bin windows in LDS (B0 / B1)
  long long img_off, sino_off;
};

// Device-side work queue of a slice batch: slices in processing order and the
// cumulative item counts (A: 64-bin waves, A^T: 16x16 pixel tiles).
struct BatchPlan {
  const SliceDesc* slices;
  const int* fwd_order;
  const long long* fwd_prefix;  // [nslice + 1]
  const int* bwd_order;
  const long long* bwd_prefix;  // [nslice + 1]
  unsigned int* counters;       // [2] queue heads (A, A^T)
  int nslice;
  long long nfwd, nbwd;
};

// Workgroup -> work item (x, y, z) of the grid, walked x fastest, such that
// the workgroups one XCD receives take a contiguous range of items.  The
// dispatcher deals the workgroups of a launch to the 8 XCDs in turn
// (workgroup n in dispatch order to XCD n mod 8); in the gather kernels z
// selects a view group, whose bin records and sinogram rows every x-y tile of
// that z reads, so each XCD's L2 then holds the records of its own groups
// instead of all of them.  Any grid: XCD k takes items [k q + min(k, r), ...)
// with q = n / 8, r = n % 8.  Used where it measured faster (SF and H0 A^T).
constexpr unsigned kXcds = 8;
__device__ __forceinline__ uint3 xcd_block() {
  const unsigned tx = gridDim.x, ty = gridDim.y, n = tx * ty * gridDim.z;
  const unsigned lin = blockIdx.x + tx * (blockIdx.y + ty * blockIdx.z);
  const unsigned k = lin % kXcds, q = n / kXcds, r = n % kXcds;
  const unsigned item = k * q + (k < r ? k : r) + lin / kXcds;
  return make_uint3(item % tx, (item / tx) % ty, item / (tx * ty));
}

constexpr int kMaxWindow = 128;   // bins of one view cached in LDS per pixel tile
constexpr int kViewChunk = 4;     // views per LDS refill in the gather kernels
constexpr int kTile = 16;         // pixel tile edge of the gather kernels

}  // namespace sgdct
