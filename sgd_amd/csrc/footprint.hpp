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
// which is the b -> 0 limit of the same expression.
//
// The kernels work in the offset t' = s t scaled to the support, s = 1/(a+b)
// (a' = s a, b' = s b, a' + b' = 1), and in units of c' = 1/(24 a'^2 b'^2):
// with z = |t'| and the three truncations, each in [0, 1],
//   y = (1 - z)_+,  m = (a' - z)_+ = (y - b')_+,  e = (b' - z)_+ = (y - a')_+,
//   S(t) / c' = sign(t') * (H(z) + 2 e^4)
//   H = h - k2 m^2 - k4           for m >= b' (y >= 2b')
//     = h - y^4 + 2 m^4           otherwise
// h = 12 a'^2 b'^2 (= 1 / (2 c')), k2 = 12 b'^2, k4 = 2 b'^4, and
//   conv = (S(tb) - S(ta)) / (tb - ta) = s c' (S/c'(tb') - S/c'(ta')) / (tb' - ta').
// Because every truncated argument lies in [0, 1], each truncation is one
// subtraction with the VALU's clamp-to-[0, 1] output modifier.  The
// tri_a * box limit (b' = 0, s = 1/a, a' = 1) has c' = infinity: its record
// holds the unscaled S instead (h = 1/2, k2 = 1/2, k4 = 0, s c' replaced by s;
// there m = y >= b' always and e = 0, so S = sign(t') (1/2 - m^2 / 2)).
// b1_mean returns the quotient without the factor s c' (`scale`); A applies it
// once per bin, A^T stages it into the sinogram value.
struct B1Bin {
  float ka, kb;      // ta' = ka * N_u(pixel), tb' = kb * N_{u+1}(pixel)
  float h, b;        // h, b'
  float a, nk2;      // a', -k2
  float hk4;         // h - k4
  float scale;       // s c' (A^T's LDS copy holds sinogram value * s c' here)
};

// Clamp to [0, 1]; folds into the clamp modifier of the producing instruction.
__device__ __forceinline__ float sat01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

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

// S(t) / c' in the scaled offset (above).
__device__ __forceinline__ float s_b1(float t, const B1Bin& r) {
  const float z = __builtin_fabsf(t);
  const float y = sat01(1.0f - z);
  const float m = sat01(r.a - z);
  const float e = sat01(r.b - z);
  const float m2 = m * m;
  const float y2 = y * y;
  const float ha = __builtin_fmaf(r.nk2, m2, r.hk4);                          // h - k2 m^2 - k4
  const float hb = __builtin_fmaf(2.0f, m2 * m2, __builtin_fmaf(-y2, y2, r.h));  // h - y^4 + 2 m^4
  const float hs = (m >= r.b) ? ha : hb;
  const float e2 = e * e;
  return copysign_f(__builtin_fmaf(2.0f, e2 * e2, hs), t);
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

// B1 footprint of the scaled interval [ta', tb'] without the factor scale.
__device__ __forceinline__ float b1_mean(const B1Bin& r, float ta, float tb) {
  return mean_over(s_b1(ta, r), s_b1(tb, r), ta, tb);
}

}  // namespace sgdct
