// sf2d_dev.hpp -- device code of the 2-D separable-footprint projector,
// shared by sf2d.hip and the batched persistent kernels (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "projector.hpp"

namespace sgdct {

struct Trap {
  float q0, q1, q2, q3, ir, iff;  // sorted corners, 1/(2(q1-q0)), 1/(2(q3-q2))
};

// min / max of finite values as v_med3_f32 against -inf / +inf: exact, one
// instruction, and no NaN quieting of the operands (fminf / fmaxf add it)
__device__ __forceinline__ float min2f(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, -__builtin_inff()); }
__device__ __forceinline__ float max2f(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, __builtin_inff()); }

__device__ __forceinline__ void sort4(float& a, float& b, float& c, float& d) {
  float t;
  t = min2f(a, b); b = max2f(a, b); a = t;
  t = min2f(c, d); d = max2f(c, d); c = t;
  t = min2f(a, c); c = max2f(a, c); a = t;
  t = min2f(b, d); d = max2f(b, d); b = t;
  t = min2f(b, c); c = max2f(b, c); b = t;
}

__device__ __forceinline__ Trap make_trap(float q0, float q1, float q2, float q3) {
  sort4(q0, q1, q2, q3);
  Trap t;
  t.q0 = q0; t.q1 = q1; t.q2 = q2; t.q3 = q3;
  t.ir = __builtin_amdgcn_rcpf(2.0f * (q1 - q0));
  t.iff = __builtin_amdgcn_rcpf(2.0f * (q3 - q2));
  return t;
}

// Area of the unit-height trapezoid over [s1, s2] (layers/A_SF.cu:48-84).
__device__ __forceinline__ float trap_gamma(const Trap& t, float s1, float s2) {
  float b1 = fmaxf(s1, t.q0), b2 = fminf(s2, t.q1);
  float d0 = b2 - t.q0, d1 = b1 - t.q0;
  float g = (b2 > b1) ? (d0 * d0 - d1 * d1) * t.ir : 0.0f;
  b1 = fmaxf(s1, t.q1);
  b2 = fminf(s2, t.q2);
  g += fmaxf(b2 - b1, 0.0f);
  b1 = fmaxf(s1, t.q2);
  b2 = fminf(s2, t.q3);
  float e1 = b1 - t.q3, e2 = b2 - t.q3;
  g += (b2 > b1) ? (e1 * e1 - e2 * e2) * t.iff : 0.0f;
  return g;
}

// dx / max(|cos t|, |sin t|) of the ray through (wx, wy) from the source;
// v_sqrt_f32 (1 ulp) instead of the correctly rounded square-root sequence
__device__ __forceinline__ float sf_weight(float dx, float wx, float wy) {
  float r2 = wx * wx + wy * wy;
  return dx * __builtin_amdgcn_sqrtf(r2) * __builtin_amdgcn_rcpf(fmaxf(fabsf(wx), fabsf(wy)));
}

// A: bin-driven gather (same strip walk as fwd_bin_kernel, corner rule h = 1/2).
// One wave: bins [64 chunk, 64 chunk + 64) of view v, wave = chunk * np + v
// (chunk-major: the four waves of a workgroup walk the same bins of adjacent
// views, nearly the same pixels; SF A 1.505 -> 1.468 ms at 512 x 512, 720
// views, 768 bins against view-major order).
// nseg > 1: lines [seg L, seg L + L) only, FP64 partial sum to part[v nu + u].
__device__ __forceinline__ void sf_fwd_wave(const Proj2DArgs& g, const float* __restrict__ img,
                                            float* __restrict__ sino, int wave, int seg = 0, int nseg = 1,
                                            double* __restrict__ part = nullptr) {
  const int lane = threadIdx.x & 63;
  const int chunk = wave / g.np;
  if (chunk >= g.nchunk) return;
  const int v = wave - chunk * g.np;
  const int u = chunk * 64 + lane;
  const ViewGeo vg = g.views[v];
  const int umid = min(chunk * 64 + 32, g.nu - 1);
  const double mx = vg.e0x + (umid + 0.5) * vg.pux, my = vg.e0y + (umid + 0.5) * vg.puy;
  const bool rows = fabs(my) >= fabs(mx);
  if (u >= g.nu) return;

  const double eax = vg.e0x + u * vg.pux, eay = vg.e0y + u * vg.puy;
  const double ebx = eax + vg.pux, eby = eay + vg.puy;
  const double acx = -eay * g.dx, acy = eax * g.dy, ac0 = eax * vg.w0y - eay * vg.w0x;
  const double bcx = -eby * g.dx, bcy = ebx * g.dy, bc0 = ebx * vg.w0y - eby * vg.w0x;
  const double ra = 0.5 * (fabs(acx) + fabs(acy)), rb = 0.5 * (fabs(bcx) + fabs(bcy));
  const double sa = rows ? acx : acy, ma = rows ? acy : acx;
  const double sb = rows ? bcx : bcy, mb = rows ? bcy : bcx;
  const int nmin = rows ? g.nx : g.ny, nmaj = rows ? g.ny : g.nx;
  const size_t smin = rows ? 1 : g.nx, smaj = rows ? g.nx : 1;
  const double isa = sa != 0.0 ? 1.0 / sa : 0.0, isb = sb != 0.0 ? 1.0 / sb : 0.0;
  // beta(ix, iy) = cross(pu, w): affine, gradient (gx1, gy1) per index step
  const double bcxx = -vg.puy * g.dx, bcyy = vg.pux * g.dy, bc00 = vg.pux * vg.w0y - vg.puy * vg.w0x;
  const float dN = (float)sa, dB = (float)(rows ? bcxx : bcyy);
  const float dwx = rows ? (float)g.dx : 0.0f, dwy = rows ? 0.0f : (float)g.dy;
  // corner offsets (+-1/2, +-1/2) of N_u and beta
  const float nx_ = 0.5f * (float)acx, ny_ = 0.5f * (float)acy;
  const float bx_ = 0.5f * (float)bcxx, by_ = 0.5f * (float)bcyy;
  const float fdx = (float)g.dx;
  // corner offsets along the walk (w) and across it (c): the corners on the
  // leading side (+w) of one pixel are the trailing side (-w) of the next
  const float w_n = rows ? nx_ : ny_, c_n = rows ? ny_ : nx_;
  const float w_b = rows ? bx_ : by_, c_b = rows ? by_ : bx_;

  int m_lo = 0, m_hi = nmaj - 1;
  if (ma != 0.0 && mb != 0.0) {
    const double last = nmin - 1;
    double c1 = (-ra - ac0) / ma, c2 = (-ra - ac0 - sa * last) / ma;
    double c3 = (rb - bc0) / mb, c4 = (rb - bc0 - sb * last) / mb;
    double lo = fmin(fmin(c1, c2), fmin(c3, c4)), hi = fmax(fmax(c1, c2), fmax(c3, c4));
    m_lo = (int)fmax(floor(lo) - 1.0, 0.0);
    m_hi = (int)fmin(ceil(hi) + 1.0, (double)(nmaj - 1));
  }
  if (nseg > 1) {
    const int len = (nmaj + nseg - 1) / nseg;
    m_lo = max(m_lo, seg * len);
    m_hi = min(m_hi, seg * len + len - 1);
  }
  double acc = 0.0;
  for (int m = m_lo; m <= m_hi; ++m) {
    const double r1 = -ra - ma * m - ac0;
    const double r2 = rb - mb * m - bc0;
    double lo = 0.0, hi = (double)(nmin - 1);
    if (sa > 0.0) lo = fmax(lo, ceil(r1 * isa));
    else if (sa < 0.0) hi = fmin(hi, floor(r1 * isa));
    else if (r1 > 0.0) continue;
    if (sb > 0.0) hi = fmin(hi, ceil(r2 * isb) - 1.0);
    else if (sb < 0.0) lo = fmax(lo, floor(r2 * isb) + 1.0);
    else if (r2 <= 0.0) continue;
    if (lo > hi) continue;
    const int n0 = (int)lo, n1 = (int)hi;
    const int ixs = rows ? n0 : m, iys = rows ? m : n0;
    float nu_ = (float)(acx * ixs + acy * iys + ac0);
    float be = (float)(bcxx * ixs + bcyy * iys + bc00);
    float wx = (float)(ixs * g.dx + vg.w0x), wy = (float)(iys * g.dy + vg.w0y);
    const float* p = img + (size_t)m * smaj + (size_t)n0 * smin;
    float accl = 0.0f;
    // corner projections relative to edge u, -N_u(c) / beta(c), of the
    // trailing corners of the line's first pixel
    float t1 = -(nu_ - w_n - c_n) * __builtin_amdgcn_rcpf(be - w_b - c_b);
    float t2 = -(nu_ - w_n + c_n) * __builtin_amdgcn_rcpf(be - w_b + c_b);
    for (int n = n0; n <= n1; ++n) {
      const float l1 = -(nu_ + w_n - c_n) * __builtin_amdgcn_rcpf(be + w_b - c_b);
      const float l2 = -(nu_ + w_n + c_n) * __builtin_amdgcn_rcpf(be + w_b + c_b);
      const Trap t = make_trap(t1, t2, l1, l2);
      t1 = l1;
      t2 = l2;
      const float gm = trap_gamma(t, 0.0f, 1.0f);
      const float val = (gm > 1e-7f) ? gm * sf_weight(fdx, wx, wy) : 0.0f;
      accl = __builtin_fmaf(val, *p, accl);
      p += smin;
      nu_ += dN;
      be += dB;
      wx += dwx;
      wy += dwy;
    }
    acc += (double)accl;
  }
  if (part) part[(size_t)v * g.nu + u] = acc;
  else sino[(size_t)v * g.nu + u] = (float)acc;
}


// A^T: pixel-driven gather, one 16x16 tile per 256-thread block, views
// [v_begin, v_end); with `part` the FP64 sums go to part[iy * nx + ix].
__device__ __forceinline__ void sf_bwd_tile(const Proj2DArgs& g, const float* __restrict__ sino,
                                            float* __restrict__ img, int tile_x, int tile_y, int v_begin = 0,
                                            int v_end = -1, double* __restrict__ part = nullptr) {
  const int ix = tile_x * kTile + (threadIdx.x & (kTile - 1));
  const int iy = tile_y * kTile + threadIdx.x / kTile;
  if (ix >= g.nx || iy >= g.ny) return;
  if (v_end < 0) v_end = g.np;
  const float fdx = (float)g.dx;
  double acc = 0.0;
  for (int v = v_begin; v < v_end; ++v) {
    const ViewGeo& vg = g.views[v];
    const double wx = ix * g.dx + vg.w0x, wy = iy * g.dy + vg.w0y;
    const double n0 = vg.e0x * wy - vg.e0y * wx;
    const double be = vg.pux * wy - vg.puy * wx;
    const double kc = -n0 / be;
    const double kfl = floor(kc);
    const float kf = (float)(kc - kfl);
    const float kcf = (float)kc, bef = (float)be;
    const float gx = 0.5f * __builtin_fmaf(kcf, vg.gx1, vg.gx0), gy = 0.5f * __builtin_fmaf(kcf, vg.gy1, vg.gy0);
    const float bx = 0.5f * vg.gx1, by = 0.5f * vg.gy1;
    // corner c: k*(c) - kc = -(d.G) / (beta + d.grad beta)
    const float q0 = kf + (gx + gy) * __builtin_amdgcn_rcpf(bef - bx - by);
    const float q1 = kf + (gx - gy) * __builtin_amdgcn_rcpf(bef - bx + by);
    const float q2 = kf - (gx - gy) * __builtin_amdgcn_rcpf(bef + bx - by);
    const float q3 = kf - (gx + gy) * __builtin_amdgcn_rcpf(bef + bx + by);
    const Trap t = make_trap(q0, q1, q2, q3);
    const int ki = (int)kfl;
    const int jlo = (int)floorf(t.q0), jhi = (int)floorf(t.q3);
    if (ki + jhi < 0 || ki + jlo > g.nu - 1) continue;
    const int ulo = max(ki + jlo, 0), uhi = min(ki + jhi, g.nu - 1);
    const float w = sf_weight(fdx, (float)wx, (float)wy);
    const float* row = sino + (size_t)v * g.nu;
    float accv = 0.0f;
    for (int u = ulo; u <= uhi; ++u) {
      const float s1 = (float)(u - ki);
      const float gm = trap_gamma(t, s1, s1 + 1.0f);
      accv = __builtin_fmaf((gm > 1e-7f) ? gm : 0.0f, row[u], accv);
    }
    acc += (double)(accv * w);
  }
  if (part) part[(size_t)iy * g.nx + ix] = acc;
  else img[(size_t)iy * g.nx + ix] = (float)acc;
}


}  // namespace sgdct
