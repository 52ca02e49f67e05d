// footprint.hpp -- 1-D box-spline footprints evaluated on the GPU (FP32).
//
// Every 2-D fan-beam basis here reduces, for one (pixel, bin) pair, to the
// average over an interval [ta, tb] of a unit-integral 1-D kernel K:
//     conv = (S(tb) - S(ta)) / (tb - ta),   S the odd antiderivative of K,
// where ta < tb are the signed offsets, from the pixel centre, at which the two
// edge rays of the bin cross the line through the pixel centre perpendicular to
// the bin's central ray.  This is exactly the quantity of
//   layers/A_B1.cu:182-197 + spline_conv.hh:372-390 (fast_2tri_1box_spline / |x2-x1|)
//   layers/A_B0.cu:211-219 + spline_conv.hh:244-293 (3-direction box spline)
// written in a form without the reference's catastrophic cancellations.
#pragma once
#include <hip/hip_runtime.h>

namespace sgdct {

// Per-bin constants of the B1 kernel K = tri_a * tri_b (unit-integral hats
// of half-widths a >= b; layers/A_B1.cu:462-463, 197).  With y = a + b - |t|:
//   S(t) = sign(t) * (1/2 - c*T(y) + 2c*max(y - a, 0)^4),  c = 1/(24 a^2 b^2)
//   T(y) = y^4 - 2(y-b)_+^4 + (y-2b)_+^4
//        = 12 b^2 (y-b)^2 + 2 b^4           for y >= 2b (expanded: no cancellation)
// For b <= 1e-7 the reference switches to tri_a * box (spline_conv.hh:393-405),
// which is the b -> 0 limit of the same expression (c terms dropped).
//
// The kernels work in the scaled offset t' = s t with s = c^(1/4) (s = 1/a in
// the tri_a * box limit), which takes c out of every evaluation: with
// y = a' + b' - |t'|, a' = s a, b' = s b,
//   S(t) = sign(t') * (1/2 - Q(y) + 2 max(y - a', 0)^4)
//   Q(y) = y^4 - 2 max(y - b', 0)^4              for y < 2b'
//        = k2 (y - b')^2 + k4                    for y >= 2b'
// k2 = 12 b'^2, k4 = 2 b'^4 (limit: b' = 0, k2 = 1/2, k4 = 0; there y <= a',
// so the last term of S vanishes), and
//   conv = (S(tb) - S(ta)) / (tb - ta) = s (S(tb') - S(ta')) / (tb' - ta').
// Traits<B1>::conv returns the quotient without the factor s; A applies
// `scale` once per bin, A^T stages it into the sinogram value.
// The record holds -k2 and 1/2 - k4, so that 1/2 - Q comes out of one FMA.
struct B1Bin {
  float ka, kb;      // ta' = ka * N_u(pixel), tb' = kb * N_{u+1}(pixel)
  float apb, b, a;   // a' + b', b', a'
  float nk2, h4;     // -k2, 1/2 - k4
  float scale;       // s (A^T's LDS copy holds sinogram value * s here)
};

// Per-bin constants of the B0 kernel K = box_p * box_q (trapezoid, p >= q;
// p, q = |vx|, |vy| of layers/A_B0.cu:369-373).  S(t) for x = |t|:
//   x <= (p-q)/2 :  x / p
//   otherwise    :  1/2 - max((p+q)/2 - x, 0)^2 / (2 p q)
struct B0Bin {
  float ka, kb;
  float hpq, hmq;    // (p+q)/2, (p-q)/2
  float inv_p, inv_2pq;
  float y;           // A^T's LDS copy: sinogram value
  float pad;
};

__device__ __forceinline__ float copysign_f(float mag, float sgn) {
  return __builtin_copysignf(mag, sgn);
}

// S(t) in the scaled offset.  With m = max(y - b', 0): y >= 2b' exactly when
// m >= b' (b' >= 0), and there m = y - b', so both branches of 1/2 - Q use m^2.
__device__ __forceinline__ float s_b1(float t, float apb, float b, float a, float nk2, float h4) {
  float x = __builtin_fabsf(t);
  float y = __builtin_fmaxf(apb - x, 0.0f);
  float m = __builtin_fmaxf(y - b, 0.0f);
  float m2 = m * m;
  float y2 = y * y;
  float ha = __builtin_fmaf(nk2, m2, h4);                                       // 1/2 - k2 m^2 - k4
  float hb = __builtin_fmaf(2.0f, m2 * m2, __builtin_fmaf(-y2, y2, 0.5f));      // 1/2 - y^4 + 2 m^4
  float h = (m >= b) ? ha : hb;
  float e = __builtin_fmaxf(y - a, 0.0f);
  float e2 = e * e;
  float s = __builtin_fmaf(2.0f, e2 * e2, h);
  return copysign_f(s, t);
}

__device__ __forceinline__ float s_b0(float t, float hpq, float hmq, float inv_p, float inv_2pq) {
  float x = __builtin_fabsf(t);
  float lo = x * inv_p;
  float r = __builtin_fmaxf(hpq - x, 0.0f);
  float hi = __builtin_fmaf(-r * r, inv_2pq, 0.5f);
  return copysign_f((x <= hmq) ? lo : hi, t);
}

// Average of K over [ta, tb] (ta < tb).  The reference also drops values
// <= 1e-7 (layers/A_B1.cu:203, A_B0.cu:225); at the bench workload that
// skip changes A by 5e-10 (B1) and 3e-11 (B0) relative L2, so it is omitted.
__device__ __forceinline__ float mean_over(float sa, float sb, float ta, float tb) {
  return (sb - sa) * __builtin_amdgcn_rcpf(tb - ta);
}

}  // namespace sgdct
