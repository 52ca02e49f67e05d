// cone3d.hip -- gfx950 kernels of the cone-beam projectors: B1 (layers/A_B1_3D.cu)
// and separable footprints SF-TR / SF-TT (layers/SF_projector.cu).
//
// The reference's value of voxel (ix, iy, iz) in detector cell (u, j) of view p
// factors into an in-plane part and an axial part (A_B1_3D.cu:95-305):
//     conv_p(u; ix, iy)  *  x * cfac_p(ix, iy, iz) * gamma_p(j; ix, iy, iz)
// conv is the 2-D B1 footprint of the (ix, iy) column (it does not depend on
// iz: the detector's u axis is in-plane), cfac = dx dy dz * path-length weight
// * inverse area of the voxel's axial trapezoid, gamma its overlap with row j.
// So per view A = (sum over columns of conv * Y), with
//     Y_p[c][j] = sum_iz x[iz][c] * cfac * gamma_j      (column c = iy*nx + ix)
// and each Y value is shared by the ~4-5 bins that see the column.  A runs as
//   axial_fwd_kernel   Y for a chunk of views (wave per column and view:
//                      lanes stage every voxel's trapezoid in LDS, then
//                      gather by detector row; deterministic)
//   conv_fwd_kernel    proj[p][j][u] = sum_c conv(u, c) Y[c][j] (wave per
//                      bin: the bin's strip walk of the 2-D A kernel, lanes
//                      over rows, no atomics)
// and A^T, which the reference does not provide (its back-projector multiplies
// by the zeroed output volume, A_B1_3D.cu:408, 436), as the exact transpose:
//   conv_bwd_kernel    Yb[c][j] = sum_u conv(u, c) proj[p][j][u]
//   axial_bwd_kernel   x[iz][c] += cfac * sum_j gamma_j Yb[c][j]
// Volumes are kept column-major (xT[c][iz]) inside, transposed on the way in
// and out, so a wave's voxel reads are contiguous.
//
// The separable-footprint projector (layers/SF_projector.cu:81-234) has the
// same structure: its u trapezoid (corners ix +- 1/2, iy +- 1/2) does not
// depend on iz, and its amplitude factors into an in-plane part
// (2 / trapezoid width) and an axial part (lsd^2 dx dy dz / (du dv r^2) times
// the inverse v-footprint area).  The kernels are templated on the model:
//   CONE_B1     B1 in-plane, axial trapezoid from x +- 1, rows [j, j+1),
//               gamma <= 1e-7 skipped (A_B1_3D.cu:300)
//   CONE_SF_TR  SF trapezoid in-plane, axial rectangle of the centre column,
//               cells centred on integers ([j - 1/2, j + 1/2], :209-216)
//   CONE_SF_TT  as SF_TR with an axial trapezoid from x +- 1/2 at iy
// For SF the view table is built with edges at half-integer detector
// coordinates, so "bin u" of the in-plane machinery is cell [u - 1/2, u + 1/2].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "footprint.hpp"
#include "projector.hpp"
#include "sf2d_dev.hpp"

namespace sgdct {

// Columns taller than kMaxNz voxels are processed in chunks of kMaxNz (the
// axial forward stages a chunk's voxel records in LDS, the axial transpose
// holds a chunk's voxel sums in registers); detector rows are staged for the
// axial transpose in windows of kMaxNv.  Neither bounds nz or nv.
constexpr int kMaxNz = 512;
constexpr int kMaxNv = 512;
// The in-plane kernels give each wave RCW chunks of 64 detector rows (a grid
// dimension runs over the chunk groups): with RCW < nv / 64 more waves are in
// flight on the row profiles, at the price of recomputing the in-plane weight
// once per group.  RCW = kAllRows keeps every row in one wave (nv <= 512).
constexpr int kAllRows = 8;
constexpr float kGammaEps = 1e-7f; // A_B1_3D.cu:300 skips gamma <= eps

enum ConeModel : int { CONE_B1 = 0, CONE_SF_TR = 1, CONE_SF_TT = 2 };

// Axial footprint of voxel iz in a column: sorted rows of its corners (B1:
// x index +- 1 at the column's iy, A_B1_3D.cu:133-137, 188-202; SF-TT:
// x +- 1/2, SF_projector.cu:136-141, 160-168; SF-TR: the centre column, a
// rectangle stored as v0 = v1, v2 = v3, :151-158) and the axial factor cf
// (B1: dx dy dz * path-length weight * 2 / widths, A_B1_3D.cu:209-210;
// SF: lsd^2 dx dy dz / (du dv r^2) / area, SF_projector.cu:193-198, 364).
struct AxialTrap {
  float v0, v1, v2, v3, cf;
  float i01, i23;  // 1 / (2 (v1 - v0)), 1 / (2 (v3 - v2)); 0 on a flat side
};

// v_med3_f32: clamp of s into [lo, hi] (lo <= hi) in one instruction, with no
// NaN quieting of the operands (every value here is finite)
__device__ __forceinline__ float clampf(float s, float lo, float hi) { return __builtin_amdgcn_fmed3f(s, lo, hi); }

__device__ __forceinline__ float inv_or_zero(float w) { return w > 0.0f ? 0.5f * __builtin_amdgcn_rcpf(w) : 0.0f; }

// 1 / d for a normal double d of float range: the float reciprocal (relative
// error ~6e-8) refined by one Newton step (error ~4e-15), in place of the
// double division's longer sequence
__device__ __forceinline__ double recip(double d) {
  const double r0 = (double)__builtin_amdgcn_rcpf((float)d);
  return __builtin_fma(r0, __builtin_fma(-d, r0, 1.0), r0);
}

struct ColumnAxial {
  float ka, kb;        // ldz / depth at the two x corners (SF-TR: both the centre)
  float ka_lo, kb_lo;  // their FP32 remainders: ka + ka_lo is the FP64 value
  float r2;            // |in-plane source -> centre|^2
};

template <int M>
__device__ __forceinline__ ColumnAxial column_axial(const Cone3DArgs& a, const ViewGeo& vg, int ix, int iy) {
  const double off = (M == CONE_B1) ? 1.0 : (M == CONE_SF_TT ? 0.5 : 0.0);
  const double wy = iy * a.g.dy + vg.w0y;
  const double wxa = (ix - off) * a.g.dx + vg.w0x, wxb = (ix + off) * a.g.dx + vg.w0x;
  const double wx = ix * a.g.dx + vg.w0x;
  // depth along the central axis = -cross(pu, X - p) / |pu|, so
  // ldz / depth = ldz |pu| / cross(p - X, pu) with |pu| = du
  const double num = a.ldz * a.du;
  const double ka = num * recip(vg.puy * wxa - vg.pux * wy);
  const double kb = num * recip(vg.puy * wxb - vg.pux * wy);
  ColumnAxial c;
  c.ka = (float)ka;
  c.kb = (float)kb;
  c.ka_lo = (float)(ka - (double)c.ka);
  c.kb_lo = (float)(kb - (double)c.kb);
  c.r2 = (float)(wx * wx + wy * wy);
  return c;
}

// Trapezoid of voxel zc (index offset from the column's middle) from the rows
// of its corners: a1, a2 at x corner a and z -+ 1/2, b1, b2 at corner b.
template <int M>
__device__ __forceinline__ AxialTrap trap_from_corners(const Cone3DArgs& a, const ColumnAxial& c, float zc, float a1,
                                                       float a2, float b1, float b2) {
  AxialTrap t;
  t.v0 = fminf(a1, b1);
  t.v3 = fmaxf(a2, b2);
  const float h1 = fmaxf(a1, b1), l2 = fminf(a2, b2);
  t.v1 = fminf(h1, l2);
  t.v2 = fmaxf(h1, l2);
  const float oz = zc * a.dz;
  if (M == CONE_B1) {
    const float w = __builtin_sqrtf(__builtin_fmaf(oz * oz, 1.0f / c.r2, 1.0f));
    t.cf = a.cvol * w * 2.0f * __builtin_amdgcn_rcpf((t.v3 - t.v0) + (t.v2 - t.v1));
  } else {
    if (M == CONE_SF_TR) {  // a1 == b1, a2 == b2: a rectangle
      t.v1 = t.v0;
      t.v2 = t.v3;
    }
    const float area = (M == CONE_SF_TR) ? (t.v3 - t.v0) : 0.5f * ((t.v3 - t.v0) + (t.v2 - t.v1));
    t.cf = a.fac * __builtin_amdgcn_rcpf((c.r2 + oz * oz) * area);
  }
  t.i01 = inv_or_zero(t.v1 - t.v0);
  t.i23 = inv_or_zero(t.v3 - t.v2);
  return t;
}

// Corners as detector rows (A's staged records).
template <int M>
__device__ __forceinline__ AxialTrap voxel_trap(const Cone3DArgs& a, const ColumnAxial& c, int iz) {
  const float zc = (float)iz - 0.5f * (float)(a.nz - 1);
  const float a1 = __builtin_fmaf(c.ka, zc - 0.5f, a.vc), a2 = __builtin_fmaf(c.ka, zc + 0.5f, a.vc);
  const float b1 = __builtin_fmaf(c.kb, zc - 0.5f, a.vc), b2 = __builtin_fmaf(c.kb, zc + 0.5f, a.vc);
  return trap_from_corners<M>(a, c, zc, a1, a2, b1, b2);
}

// The same trapezoid with its corners relative to the integer row *k0 next
// to it.  As FP32 detector rows in the hundreds, corners carry up to ~6e-5
// rows of rounding, which A^T of white-noise rows reads (~8e-6 relative at
// 512^3 voxels / 768 rows); relative to k0, each corner is one fused
// multiply-add of the exact vc - k0 plus the slope's FP32 remainder, ~1e-7
// rows from exact.
template <int M>
__device__ __forceinline__ AxialTrap voxel_trap_rel(const Cone3DArgs& a, const ColumnAxial& c, int iz, int* k0) {
  const float zc = (float)iz - 0.5f * (float)(a.nz - 1);
  const float zl = zc - 0.5f;
  const int k = (int)floorf(__builtin_fmaf(c.ka, zl, a.vc));
  const float vk = a.vc - (float)k;  // exact: both are multiples of 1/2 below 2^23
  const float a1 = __builtin_fmaf(c.ka_lo, zl, __builtin_fmaf(c.ka, zl, vk));
  const float b1 = __builtin_fmaf(c.kb_lo, zl, __builtin_fmaf(c.kb, zl, vk));
  // the far corners are one slope further (zc + 1/2 = zl + 1)
  const float a2 = a1 + c.ka, b2 = b1 + c.kb;
  *k0 = k;
  return trap_from_corners<M>(a, c, zc, a1, a2, b1, b2);
}

// detector row j spans [j + row_off, j + row_off + 1)
template <int M>
__device__ __forceinline__ float row_off() {
  return M == CONE_B1 ? 0.0f : -0.5f;
}

// Overlap of [s, s+1) with the unit-height trapezoid (A_SF.cu:48-84 /
// A_B1_3D.cu:30-66), segment by segment: [s, s+1) clamped into each segment
// gives the overlap's ends (equal when they do not meet).  The sloped sides
// divide by their widths through the voxel's precomputed reciprocals, zero
// for a side of zero width, which never overlaps.
__device__ __forceinline__ float trap_gamma(float s, const AxialTrap& t) {
  const float e = s + 1.0f;
  const float s0 = clampf(s, t.v0, t.v1), e0 = clampf(e, t.v0, t.v1);
  const float s1 = clampf(s, t.v1, t.v2), e1 = clampf(e, t.v1, t.v2);
  const float s2 = clampf(s, t.v2, t.v3), e2 = clampf(e, t.v2, t.v3);
  const float g0 = (e0 - s0) * (e0 + s0 - 2.0f * t.v0) * t.i01;
  const float g2 = (e2 - s2) * (2.0f * t.v3 - s2 - e2) * t.i23;
  return g0 + (e1 - s1) + g2;
}

template <int M>
__device__ __forceinline__ bool keep_gamma(float gm) {
  return M == CONE_B1 ? gm > kGammaEps : gm == gm;
}

// Orders a wave's LDS accesses across lanes: LDS operations of one wave are
// performed in program order, so a compiler barrier at wavefront scope suffices.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Y[pl][c][j] = sum_iz xT[c][iz] * cfac * gamma_j.  Block: 4 waves, one
// column each (each wave uses only its own LDS records); grid
// (ceil(ncol / 4), views in the chunk).  BIG (nz > kMaxNz): the column runs
// in chunks of kMaxNz voxels, the rows' sums added chunk by chunk.
// ---------------------------------------------------------------------------
// Dynamic LDS: per wave and voxel a record of kAxialFields floats (v0, v1,
// v2, v3, i01, i23, x * cf).  The odd stride puts the records of 64
// consecutive voxels on distinct banks, so a wave's rows (which read
// neighbouring voxels) do not conflict.
constexpr int kAxialFields = 7;

template <int M, bool BIG>
__global__ __launch_bounds__(256) void axial_fwd_kernel(Cone3DArgs a, int p0, const float* __restrict__ xT,
                                                        float* __restrict__ Y) {
  extern __shared__ float axial_lds[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int zcap = BIG ? kMaxNz : a.nz;
  float* const R = axial_lds + (size_t)wv * kAxialFields * zcap;  // this wave's records
  const int ncol = a.g.nx * a.g.ny;
  const int c = blockIdx.x * 4 + wv, pl = blockIdx.y;
  const bool live = c < ncol;
  ColumnAxial ca{};
  if (live) ca = column_axial<M>(a, a.g.views[p0 + pl], c % a.g.nx, c / a.g.nx);
  float* yrow = Y + ((size_t)pl * ncol + c) * a.nv;
  // v3(iz) = vc + k (zc + 1/2), zc = iz - (nz - 1) / 2, with k = max(ka, kb)
  // above vc and min(ka, kb) below; it grows with iz (ka, kb > 0: voxels in
  // front of the source), so the first voxel reaching a row follows from that
  // line up to rounding, and the staged values settle it exactly
  const float ikhi = 1.0f / fmaxf(ca.ka, ca.kb), iklo = 1.0f / fminf(ca.ka, ca.kb);
  const float zoff = 0.5f * (float)(a.nz - 1) - 0.5f;
  // voxels [z0, z1): stage their records, then add their share to every row
  auto chunk = [&](int z0, int z1, bool first) {
    if (live) {
      for (int iz = z0 + lane; iz < z1; iz += 64) {
        const AxialTrap t = voxel_trap<M>(a, ca, iz);
        float* const r = R + kAxialFields * (iz - z0);
        r[0] = t.v0;
        r[1] = t.v1;
        r[2] = t.v2;
        r[3] = t.v3;
        r[4] = t.i01;
        r[5] = t.i23;
        r[6] = xT[(size_t)c * a.nz + iz] * t.cf;
      }
    }
    __syncthreads();
    if (!live) return;
    for (int j = lane; j < a.nv; j += 64) {
      const float fj = (float)j + row_off<M>();
      // first voxel of the chunk whose footprint reaches row j
      const float est = __builtin_fmaf(fj - a.vc, fj >= a.vc ? ikhi : iklo, zoff);
      int lo = (int)fminf(fmaxf(ceilf(est), (float)z0), (float)z1);
      while (lo > z0 && R[kAxialFields * (lo - 1 - z0) + 3] >= fj) --lo;
      while (lo < z1 && R[kAxialFields * (lo - z0) + 3] < fj) ++lo;
      float acc = 0.0f;
      for (int iz = lo; iz < z1; ++iz) {
        const float* const r = R + kAxialFields * (iz - z0);
        if (!(r[0] < fj + 1.0f)) break;
        AxialTrap t;
        t.v0 = r[0];
        t.v1 = r[1];
        t.v2 = r[2];
        t.v3 = r[3];
        t.i01 = r[4];
        t.i23 = r[5];
        const float gm = trap_gamma(fj, t);
        if (keep_gamma<M>(gm)) acc = __builtin_fmaf(r[6], gm, acc);
      }
      if (first) yrow[j] = acc;
      else yrow[j] += acc;
    }
  };
  if (!BIG) {
    chunk(0, a.nz, true);
    return;
  }
  // every wave of the block reaches both barriers of every chunk
  for (int z0 = 0; z0 < a.nz; z0 += kMaxNz) {
    if (z0 > 0) __syncthreads();  // the previous chunk's reads come first
    chunk(z0, min(a.nz, z0 + kMaxNz), z0 == 0);
  }
}

// ---------------------------------------------------------------------------
// proj[p][j][u] = sum over the columns c of bin u's strip of conv(u, c) Y[c][j].
// Wave per (view, bin); the strip walk is the 2-D B1 A kernel's with the bin
// wave-uniform.  Grid: ceil(nu / 4) * nview blocks, view = blockIdx % nview
// (consecutive blocks, which the dispatcher spreads over the XCDs, take
// different views; the blocks of one view share an XCD's L2 when nview is 8).
// ---------------------------------------------------------------------------
// SF in-plane weight of cell u for pixel (ix, iy): area of the pixel's unit
// trapezoid over the cell times 2 / (sum of its widths), from the corner
// projections relative to edge u, -N_u(c) / beta(c) (SF_projector.cu:123-133,
// 175-190, 193-198, 209-214).
__device__ __forceinline__ float sf_cell_weight(const ViewGeo& vg, const Proj2DArgs& g, int u, int ix, int iy) {
  const double wx = ix * g.dx + vg.w0x, wy = iy * g.dy + vg.w0y;
  const double ex = vg.e0x + u * vg.pux, ey = vg.e0y + u * vg.puy;
  const float nc = (float)(ex * wy - ey * wx);
  const float bc = (float)(vg.pux * wy - vg.puy * wx);
  const float hx = (float)(-0.5 * ey * g.dx), hy = (float)(0.5 * ex * g.dy);      // N_u steps
  const float bx = (float)(-0.5 * vg.puy * g.dx), by = (float)(0.5 * vg.pux * g.dy);  // beta steps
  const float q0 = -(nc - hx - hy) / (bc - bx - by);
  const float q1 = -(nc - hx + hy) / (bc - bx + by);
  const float q2 = -(nc + hx - hy) / (bc + bx - by);
  const float q3 = -(nc + hx + hy) / (bc + bx + by);
  const Trap t = make_trap(q0, q1, q2, q3);
  const float width = (t.q3 - t.q0) + (t.q2 - t.q1);
  const float gm = trap_gamma(t, 0.0f, 1.0f);
  return width > 0.0f ? gm * 2.0f / width : 0.0f;
}

template <int M, int RCW>
__global__ __launch_bounds__(256) void conv_fwd_kernel(Cone3DArgs a, int p0, int nview, const float* __restrict__ Y,
                                                       float* __restrict__ proj) {
  const Proj2DArgs& g = a.g;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kc0 = blockIdx.y * RCW;  // this wave's detector-row chunks
  const int pl = blockIdx.x % nview;
  const int u = (blockIdx.x / nview) * 4 + wv;
  if (u >= g.nu) return;
  const int v = p0 + pl;
  const int ncol = g.nx * g.ny;
  const ViewGeo vg = g.views[v];
  B1Bin rec{};
  if (M == CONE_B1) rec = reinterpret_cast<const B1Bin*>(g.bins)[(size_t)v * g.nu + u];

  const double eax = vg.e0x + u * vg.pux, eay = vg.e0y + u * vg.puy;
  const double ebx = eax + vg.pux, eby = eay + vg.puy;
  const double mx = eax + 0.5 * vg.pux, my = eay + 0.5 * vg.puy;
  const bool rows = fabs(my) >= fabs(mx);
  const double acx = -eay * g.dx, acy = eax * g.dy, ac0 = eax * vg.w0y - eay * vg.w0x;
  const double bcx = -eby * g.dx, bcy = ebx * g.dy, bc0 = ebx * vg.w0y - eby * vg.w0x;
  // clamped end bins (A_B1_3D.cu:150-151, as in the 2-D kernel); SF drops
  // cells outside the detector instead
  const bool clamp = M == CONE_B1;
  const double lax = (clamp && u == 0) ? eax - vg.pux : eax, lay = (clamp && u == 0) ? eay - vg.puy : eay;
  const double lbx = (clamp && u == g.nu - 1) ? ebx + vg.pux : ebx;
  const double lby = (clamp && u == g.nu - 1) ? eby + vg.puy : eby;
  const double qcx = -lay * g.dx, qcy = lax * g.dy, qc0 = lax * vg.w0y - lay * vg.w0x;
  const double rcx = -lby * g.dx, rcy = lbx * g.dy, rc0 = lbx * vg.w0y - lby * vg.w0x;
  const double ra = g.h * (fabs(qcx) + fabs(qcy)), rb = g.h * (fabs(rcx) + fabs(rcy));
  const double sa = rows ? acx : acy, ma = rows ? acy : acx;
  const double sb = rows ? bcx : bcy, mb = rows ? bcy : bcx;
  const double sq = rows ? qcx : qcy, mq = rows ? qcy : qcx;
  const double sr = rows ? rcx : rcy, mr = rows ? rcy : rcx;
  const int nmin = rows ? g.nx : g.ny, nmaj = rows ? g.ny : g.nx;
  const int smin = rows ? 1 : g.nx, smaj = rows ? g.nx : 1;
  const float dta = (float)(rec.ka * sa), dtb = (float)(rec.kb * sb);

  int m_lo = 0, m_hi = nmaj - 1;
  if (mq != 0.0 && mr != 0.0) {
    const double last = nmin - 1;
    double c1 = (-ra - qc0) / mq, c2 = (-ra - qc0 - sq * last) / mq;
    double c3 = (rb - rc0) / mr, c4 = (rb - rc0 - sr * last) / mr;
    double lo = fmin(fmin(c1, c2), fmin(c3, c4)), hi = fmax(fmax(c1, c2), fmax(c3, c4));
    m_lo = (int)fmax(floor(lo) - 1.0, 0.0);
    m_hi = (int)fmin(ceil(hi) + 1.0, (double)(nmaj - 1));
  }
  double lo0 = 0.0, lo1 = 0.0, hi0 = nmin - 1, hi1 = 0.0;
  if (sq > 0.0 && sr > 0.0) {
    lo0 = (-ra - qc0) / sq; lo1 = -mq / sq; hi0 = (rb - rc0) / sr; hi1 = -mr / sr;
  } else if (sq < 0.0 && sr < 0.0) {
    hi0 = (-ra - qc0) / sq; hi1 = -mq / sq; lo0 = (rb - rc0) / sr; lo1 = -mr / sr;
  }
  const double an = (double)rec.ka * sa, am = (double)rec.ka * ma, a0 = (double)rec.ka * ac0;
  const double bn = (double)rec.kb * sb, bm = (double)rec.kb * mb, b0 = (double)rec.kb * bc0;

  double acc[RCW];
#pragma unroll
  for (int k = 0; k < RCW; ++k) acc[k] = 0.0;
  // Row-profile reads of a column: every lane loads (rows past nv read the
  // last row and are never stored), so the loads need no per-lane branch.
  const float* ybase = Y + (size_t)pl * ncol * a.nv + kc0 * 64 + lane;
  const int jmax = a.nv - 1 - kc0 * 64 - lane;  // offset of the last row
  auto load_col = [&](int m, int n, float* y) {
    const float* yc = ybase + (size_t)(m * smaj + n * smin) * a.nv;
#pragma unroll
    for (int k = 0; k < RCW; ++k)
      if ((kc0 + k) * 64 < a.nv) y[k] = yc[min(k * 64, jmax)];
  };
  // strip range of line m (empty when n0 > n1)
  auto line_range = [&](int m, int& n0, int& n1) {
    n0 = (int)fmax(ceil(fma((double)m, lo1, lo0)), 0.0);
    n1 = (int)fmin(floor(fma((double)m, hi1, hi0)), (double)(nmin - 1));
  };
  // Lines are walked in order; within a line the next column's profile is
  // loaded while the current column's weight is computed, and the last
  // column of a line loads the first of the next line.
  float ycur[RCW], ynext[RCW];
#pragma unroll
  for (int k = 0; k < RCW; ++k) ycur[k] = ynext[k] = 0.0f;
  int n0, n1;
  // m_lo is not clamped from above: a strip that misses the volume has
  // m_lo > m_hi and loads nothing
  n0 = 0;
  n1 = -1;
  if (m_lo <= m_hi) line_range(m_lo, n0, n1);
  if (n0 <= n1) load_col(m_lo, n0, ycur);
  for (int m = m_lo; m <= m_hi; ++m) {
    int q0 = 0, q1 = -1;  // next line's range
    if (m < m_hi) line_range(m + 1, q0, q1);
    if (n0 <= n1) {
      float ta = (float)fma((double)n0, an, fma((double)m, am, a0));
      float tb = (float)fma((double)n0, bn, fma((double)m, bm, b0));
      float accl[RCW];
#pragma unroll
      for (int k = 0; k < RCW; ++k) accl[k] = 0.0f;
      for (int n = n0; n <= n1; ++n) {
        if (n < n1) load_col(m, n + 1, ynext);
        else if (q0 <= q1) load_col(m + 1, q0, ynext);
        float w;
        if (M == CONE_B1) {
          w = b1_mean(rec, ta, tb) * rec.scale;
        } else {
          const int ix = rows ? n : m, iy = rows ? m : n;
          w = sf_cell_weight(vg, g, u, ix, iy);
        }
#pragma unroll
        for (int k = 0; k < RCW; ++k) {
          accl[k] = __builtin_fmaf(w, ycur[k], accl[k]);
          ycur[k] = ynext[k];
        }
        ta += dta;
        tb += dtb;
      }
#pragma unroll
      for (int k = 0; k < RCW; ++k) acc[k] += (double)accl[k];
    } else if (q0 <= q1) {
      load_col(m + 1, q0, ycur);
    }
    n0 = q0;
    n1 = q1;
  }
  float* out = proj + (size_t)v * a.nv * g.nu + u;
#pragma unroll
  for (int k = 0; k < RCW; ++k) {
    const int j = (kc0 + k) * 64 + lane;
    if (j < a.nv) out[(size_t)j * g.nu] = (float)acc[k];
  }
}

// ---------------------------------------------------------------------------
// Yb[pl][c][j] = sum_u conv(u, c) projT[pl][u][j].  Wave per (column, view),
// the column's clamped corner range (A_B1_3D.cu:117-151) wave-uniform.
// ---------------------------------------------------------------------------
template <int M, int RCW>
__global__ __launch_bounds__(256) void conv_bwd_kernel(Cone3DArgs a, int p0, const float* __restrict__ projT,
                                                       float* __restrict__ Yb) {
  const Proj2DArgs& g = a.g;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ncol = g.nx * g.ny;
  const int c = blockIdx.x * 4 + wv, pl = blockIdx.y;
  const int kc0 = blockIdx.z * RCW;  // this wave's detector-row chunks
  if (c >= ncol) return;
  const int v = p0 + pl;
  const int ix = c % g.nx, iy = c / g.nx;
  const ViewGeo& vg = g.views[v];
  const B1Bin* bins = M == CONE_B1 ? reinterpret_cast<const B1Bin*>(g.bins) + (size_t)v * g.nu : nullptr;
  const double wx = ix * g.dx + vg.w0x, wy = iy * g.dy + vg.w0y;
  const double n0 = vg.e0x * wy - vg.e0y * wx;
  const double be = vg.pux * wy - vg.puy * wx;
  const float bef = (float)be;
  const float kc = -(float)n0 * __builtin_amdgcn_rcpf(bef);
  const float gx = __builtin_fmaf(kc, vg.gx1, vg.gx0), gy = __builtin_fmaf(kc, vg.gy1, vg.gy0);
  const float hw = g.h * (__builtin_fabsf(gx) + __builtin_fabsf(gy)) * __builtin_amdgcn_rcpf(__builtin_fabsf(bef) - vg.db);
  const float last = (float)(g.nu - 1);
  const int ulo = (int)fminf(fmaxf(floorf(kc - hw), 0.0f), last);
  const int uhi = (int)fminf(fmaxf(floorf(kc + hw), 0.0f), last);
  float acc[RCW];
#pragma unroll
  for (int k = 0; k < RCW; ++k) acc[k] = 0.0f;
  // every lane loads (rows past nv read the last row and are never stored);
  // the next bin's rows are loaded while the current bin's weight is computed
  const float* rows0 = projT + (size_t)pl * g.nu * a.nv + kc0 * 64 + lane;
  const int jmax = a.nv - 1 - kc0 * 64 - lane;
  auto load_bin = [&](int u, float* y) {
    const float* row = rows0 + (size_t)u * a.nv;
#pragma unroll
    for (int k = 0; k < RCW; ++k)
      if ((kc0 + k) * 64 < a.nv) y[k] = row[min(k * 64, jmax)];
  };
  float ycur[RCW], ynext[RCW];
#pragma unroll
  for (int k = 0; k < RCW; ++k) ycur[k] = ynext[k] = 0.0f;
  load_bin(ulo, ycur);
  float na = (float)fma((double)ulo, be, n0);
  for (int u = ulo; u <= uhi; ++u) {
    if (u < uhi) load_bin(u + 1, ynext);
    const float nb = na + bef;
    float w;
    if (M == CONE_B1) {
      const B1Bin rec = bins[u];
      const float ta = rec.ka * na, tb = rec.kb * nb;
      w = b1_mean(rec, ta, tb) * rec.scale;
    } else {
      w = sf_cell_weight(vg, g, u, ix, iy);
    }
#pragma unroll
    for (int k = 0; k < RCW; ++k) {
      acc[k] = __builtin_fmaf(w, ycur[k], acc[k]);
      ycur[k] = ynext[k];
    }
    na = nb;
  }
  float* out = Yb + ((size_t)pl * ncol + c) * a.nv + kc0 * 64 + lane;
#pragma unroll
  for (int k = 0; k < RCW; ++k)
    if ((kc0 + k) * 64 + lane < a.nv) out[k * 64] = acc[k];
}

// ---------------------------------------------------------------------------
// xT[c][iz] (+)= sum over the chunk's views of cfac * sum_j gamma_j Yb[c][j].
// Wave per column, lanes over voxels; grid ceil(ncol / 4).  Each view's row
// profile Yb[c][*] is copied into the wave's LDS with coalesced loads (the
// next view's loads are in flight meanwhile); the voxels then gather their
// rows from LDS.  BIG (nz > kMaxNz or nv > kMaxNv): the grid's y dimension
// takes the column's voxels in chunks of kMaxNz, and each view's profile is
// staged kMaxNv rows at a time.
// ---------------------------------------------------------------------------
template <int M, bool BIG>
__global__ __launch_bounds__(256) void axial_bwd_kernel(Cone3DArgs a, int p0, int nview, const float* __restrict__ Yb,
                                                        float* __restrict__ xT, int accumulate) {
  __shared__ float prof_lds[4][kMaxNv];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ncol = a.g.nx * a.g.ny;
  const int c = blockIdx.x * 4 + wv;
  if (c >= ncol) return;  // waves share no LDS and meet at no block barrier
  float* const yl = prof_lds[wv];
  const int ix = c % a.g.nx, iy = c / a.g.nx;
  constexpr int kSlots = kMaxNz / 64, kRowSlots = kMaxNv / 64;
  const int z0 = BIG ? blockIdx.y * kMaxNz : 0;
  const int nwin = BIG ? (a.nv + kMaxNv - 1) / kMaxNv : 1;
  double acc[kSlots];
#pragma unroll
  for (int s = 0; s < kSlots; ++s) acc[s] = 0.0;
  float pre[kRowSlots];
  // window q = pl * nwin + w: rows [w kMaxNv, min(nv, (w + 1) kMaxNv)) of view pl
  auto fetch = [&](int q) {
    const int pl = BIG ? q / nwin : q, r0 = BIG ? (q - pl * nwin) * kMaxNv : 0;
    const float* yrow = Yb + ((size_t)pl * ncol + c) * a.nv + r0;
    const int nr = BIG ? min(kMaxNv, a.nv - r0) : a.nv;
#pragma unroll
    for (int k = 0; k < kRowSlots; ++k)
      if (k * 64 < nr) pre[k] = (k * 64 + lane < nr) ? yrow[k * 64 + lane] : 0.0f;
  };
  fetch(0);
  const float lastj = (float)(a.nv - 1);
  const int nq = nview * nwin;
  for (int q = 0; q < nq; ++q) {
    const int pl = BIG ? q / nwin : q, r0 = BIG ? (q - pl * nwin) * kMaxNv : 0;
    const int nr = BIG ? min(kMaxNv, a.nv - r0) : a.nv;
    wave_lds_sync();  // the previous window's reads come first
#pragma unroll
    for (int k = 0; k < kRowSlots; ++k)
      if (k * 64 + lane < nr) yl[k * 64 + lane] = pre[k];
    wave_lds_sync();
    if (q + 1 < nq) fetch(q + 1);
    const ColumnAxial ca = column_axial<M>(a, a.g.views[p0 + pl], ix, iy);
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
      const int iz = z0 + s * 64 + lane;
      if (iz >= a.nz) continue;
      int k0;
      const AxialTrap t = voxel_trap_rel<M>(a, ca, iz, &k0);
      const float ro = row_off<M>();
      // rows relative to k0 (small integers plus ro: exact)
      int jlo = (int)fminf(fmaxf(floorf(t.v0 - ro) + (float)k0, 0.0f), lastj);
      int jhi = (int)fminf(fmaxf(floorf(t.v3 - ro) + (float)k0, 0.0f), lastj);
      if (BIG) {
        jlo = max(jlo, r0);
        jhi = min(jhi, r0 + nr - 1);
      }
      float sum = 0.0f;
      float sj = (float)(jlo - k0) + ro;
      for (int j = jlo; j <= jhi; ++j, sj += 1.0f) {
        const float gm = trap_gamma(sj, t);
        if (keep_gamma<M>(gm)) sum = __builtin_fmaf(gm, yl[j - r0], sum);
      }
      acc[s] += (double)(t.cf * sum);
    }
  }
  float* out = xT + (size_t)c * a.nz;
#pragma unroll
  for (int s = 0; s < kSlots; ++s) {
    const int iz = z0 + s * 64 + lane;
    if (iz < a.nz) out[iz] = (float)((accumulate ? (double)out[iz] : 0.0) + acc[s]);
  }
}

// ---------------------------------------------------------------------------
// out[b][c][r] = in[b][r][c] for a batch of rows x cols matrices.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                        int rows, int cols) {
  __shared__ float tile[32][33];
  const size_t off = (size_t)blockIdx.z * rows * cols;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    if (r < rows && c < cols) tile[k][tx] = in[off + (size_t)r * cols + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + tx;
    if (r < rows && c < cols) out[off + (size_t)c * rows + r] = tile[tx][k];
  }
}

hipError_t launch_transpose(const float* in, float* out, int rows, int cols, int batch, hipStream_t s) {
  dim3 grid((cols + 31) / 32, (rows + 31) / 32, batch);
  transpose_kernel<<<grid, dim3(256), 0, s>>>(in, out, rows, cols);
  return hipGetLastError();
}

template <int M>
static void cone_forward_chunk(const Cone3DArgs& a, int p0, int nview, const float* xT, float* Y, float* proj,
                               hipStream_t s) {
  const int ncol = a.g.nx * a.g.ny;
  const size_t lds = sizeof(float) * 4 * kAxialFields * std::min(a.nz, kMaxNz);  // 4 waves
  if (a.nz > kMaxNz) axial_fwd_kernel<M, true><<<dim3((ncol + 3) / 4, nview), dim3(256), lds, s>>>(a, p0, xT, Y);
  else axial_fwd_kernel<M, false><<<dim3((ncol + 3) / 4, nview), dim3(256), lds, s>>>(a, p0, xT, Y);
  // every row in one wave: with the next column's profile loaded during the
  // weight computation, splitting the rows over more waves (each computing
  // the weights again) no longer pays (B1, 16 views of 256^2 x 384: 2 chunks
  // of 64 rows per wave 67 ms per call, 3 chunks 61 ms, all 55 ms)
  constexpr int RCW = kAllRows;
  const int groups = (a.nrc + RCW - 1) / RCW;
  conv_fwd_kernel<M, RCW><<<dim3(((a.g.nu + 3) / 4) * nview, groups), dim3(256), 0, s>>>(a, p0, nview, Y, proj);
}

template <int M>
static void cone_backward_chunk(const Cone3DArgs& a, int p0, int nview, const float* projT, float* Yb, float* xT,
                                int accumulate, hipStream_t s) {
  const int ncol = a.g.nx * a.g.ny;
  const int groups = (a.nrc + kAllRows - 1) / kAllRows;  // 512 detector rows per wave
  conv_bwd_kernel<M, kAllRows><<<dim3((ncol + 3) / 4, nview, groups), dim3(256), 0, s>>>(a, p0, projT, Yb);
  const dim3 grid((ncol + 3) / 4, (a.nz + kMaxNz - 1) / kMaxNz);
  if (a.nv > kMaxNv || a.nz > kMaxNz)
    axial_bwd_kernel<M, true><<<grid, dim3(256), 0, s>>>(a, p0, nview, Yb, xT, accumulate);
  else
    axial_bwd_kernel<M, false><<<grid, dim3(256), 0, s>>>(a, p0, nview, Yb, xT, accumulate);
}

hipError_t launch_cone_forward_chunk(int model, const Cone3DArgs& a, int p0, int nview, const float* xT, float* Y,
                                     float* proj, hipStream_t s) {
  if (model == CONE_B1) cone_forward_chunk<CONE_B1>(a, p0, nview, xT, Y, proj, s);
  else if (model == CONE_SF_TR) cone_forward_chunk<CONE_SF_TR>(a, p0, nview, xT, Y, proj, s);
  else if (model == CONE_SF_TT) cone_forward_chunk<CONE_SF_TT>(a, p0, nview, xT, Y, proj, s);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_cone_backward_chunk(int model, const Cone3DArgs& a, int p0, int nview, const float* projT,
                                      float* Yb, float* xT, int accumulate, hipStream_t s) {
  if (model == CONE_B1) cone_backward_chunk<CONE_B1>(a, p0, nview, projT, Yb, xT, accumulate, s);
  else if (model == CONE_SF_TR) cone_backward_chunk<CONE_SF_TR>(a, p0, nview, projT, Yb, xT, accumulate, s);
  else if (model == CONE_SF_TT) cone_backward_chunk<CONE_SF_TT>(a, p0, nview, projT, Yb, xT, accumulate, s);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace sgdct
