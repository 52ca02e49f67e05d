// proj2d_dev.hpp -- device code of the 2-D B0 / B1 projectors, shared by the
// per-projector kernels (proj2d.hip) and the batched persistent kernels
// (batch.hip).  See proj2d.hip for the kernel design.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "footprint.hpp"
#include "projector.hpp"

namespace sgdct {

// ---------------------------------------------------------------------------
// Basis traits: per-bin record, footprint of one (pixel, bin) pair.
// ---------------------------------------------------------------------------
template <int BASIS> struct Traits;

template <> struct Traits<BASIS_B1> {
  using Bin = B1Bin;
  // times scale(r) it is the footprint
  __device__ static __forceinline__ float conv(const Bin& r, float ta, float tb) { return b1_mean(r, ta, tb); }
  __host__ __device__ static __forceinline__ float scale(const Bin& r) { return r.scale; }
  // s of the scaled offset t' = s t: scale = s c' and h = 1 / (2 c') (the
  // tri_a * box record: scale = s, h = 1/2)
  __host__ __device__ static __forceinline__ float tscale(const Bin& r) { return 2.0f * r.h * r.scale; }
};

template <> struct Traits<BASIS_B0> {
  using Bin = B0Bin;
  __device__ static __forceinline__ float conv(const Bin& r, float ta, float tb) {
    float sa = s_b0(ta, r.hpq, r.hmq, r.inv_p, r.inv_2pq);
    float sb = s_b0(tb, r.hpq, r.hmq, r.inv_p, r.inv_2pq);
    return mean_over(sa, sb, ta, tb);
  }
  __host__ __device__ static __forceinline__ float scale(const Bin&) { return 1.0f; }
  __host__ __device__ static __forceinline__ float tscale(const Bin&) { return 1.0f; }
};

// ---------------------------------------------------------------------------
// A: bin-driven gather.
// ---------------------------------------------------------------------------
// One wave: bins [64 chunk, 64 chunk + 64) of view v, wave = v * nchunk + chunk.
// With nseg > 1 the wave covers only lines [seg L, seg L + L) of its walk
// (L = ceil(lines / nseg)) and writes its FP64 partial sum, already scaled,
// to part[v * nu + u]; otherwise it writes the finished value to sino.
template <int BASIS>
__device__ __forceinline__ void fwd_bin_wave(const Proj2DArgs& g, const float* __restrict__ img,
                                             float* __restrict__ sino, int wave, int seg = 0, int nseg = 1,
                                             double* __restrict__ part = nullptr) {
  using T = Traits<BASIS>;
  const int lane = threadIdx.x & 63;
  const int v = wave / g.nchunk;
  if (v >= g.np) return;
  const int chunk = wave - v * g.nchunk;
  const int u = chunk * 64 + lane;

  const ViewGeo vg = g.views[v];
  // Walk rows when the wave's middle ray is closer to the y axis, columns
  // otherwise (wave-uniform choice).
  const int umid = min(chunk * 64 + 32, g.nu - 1);
  const double mx = vg.e0x + (umid + 0.5) * vg.pux, my = vg.e0y + (umid + 0.5) * vg.puy;
  const bool rows = fabs(my) >= fabs(mx);
  if (u >= g.nu) return;

  const typename T::Bin rec = reinterpret_cast<const typename T::Bin*>(g.bins)[(size_t)v * g.nu + u];
  // N_u and N_{u+1} as affine functions of the pixel indices.
  const double eax = vg.e0x + u * vg.pux, eay = vg.e0y + u * vg.puy;
  const double ebx = eax + vg.pux, eby = eay + vg.puy;
  const double acx = -eay * g.dx, acy = eax * g.dy, ac0 = eax * vg.w0y - eay * vg.w0x;
  const double bcx = -eby * g.dx, bcy = ebx * g.dy, bc0 = ebx * vg.w0y - eby * vg.w0x;
  // Corner rule: bin u is in a pixel's range iff N_u + R_u >= 0 and
  // N_{u+1} - R_{u+1} < 0 (max / min of N over the pixel's corners).  The
  // reference clamps ranges into [0, nu-1] (layers/A_B1.cu:150-151), so bin 0
  // also serves pixels projecting before the detector and bin nu-1 pixels
  // projecting past it; one extra bin of margin covers every pixel whose
  // footprint can still reach the end bins.
  const double lax = (u == 0) ? eax - vg.pux : eax, lay = (u == 0) ? eay - vg.puy : eay;
  const double lbx = (u == g.nu - 1) ? ebx + vg.pux : ebx, lby = (u == g.nu - 1) ? eby + vg.puy : eby;
  const double qcx = -lay * g.dx, qcy = lax * g.dy, qc0 = lax * vg.w0y - lay * vg.w0x;
  const double rcx = -lby * g.dx, rcy = lbx * g.dy, rc0 = lbx * vg.w0y - lby * vg.w0x;
  const double ra = g.h * (fabs(qcx) + fabs(qcy)), rb = g.h * (fabs(rcx) + fabs(rcy));
  // ta / tb step along the walk (edges u, u+1); pixel-set tests (edges q, r)
  const double sa = rows ? acx : acy, ma = rows ? acy : acx;
  const double sb = rows ? bcx : bcy, mb = rows ? bcy : bcx;
  const double sq = rows ? qcx : qcy, mq = rows ? qcy : qcx;
  const double sr = rows ? rcx : rcy, mr = rows ? rcy : rcx;
  const int nmin = rows ? g.nx : g.ny, nmaj = rows ? g.ny : g.nx;
  const int smin = rows ? 1 : g.nx, smaj = rows ? g.nx : 1;
  const double isq = sq != 0.0 ? 1.0 / sq : 0.0, isr = sr != 0.0 ? 1.0 / sr : 0.0;
  const float dta = (float)(rec.ka * sa), dtb = (float)(rec.kb * sb);

  // Lines crossed by the strip: between the extreme crossings of its two
  // boundary lines with the first and last pixel of a line.
  int m_lo = 0, m_hi = nmaj - 1;
  if (mq != 0.0 && mr != 0.0) {
    const double last = nmin - 1;
    double c1 = (-ra - qc0) / mq, c2 = (-ra - qc0 - sq * last) / mq;
    double c3 = (rb - rc0) / mr, c4 = (rb - rc0 - sr * last) / mr;
    double lo = fmin(fmin(c1, c2), fmin(c3, c4)), hi = fmax(fmax(c1, c2), fmax(c3, c4));
    lo = fmax(floor(lo) - 1.0, 0.0);
    hi = fmin(ceil(hi) + 1.0, (double)(nmaj - 1));
    m_lo = (int)lo;
    m_hi = (int)hi;
  }
  if (nseg > 1) {
    const int len = (nmaj + nseg - 1) / nseg;
    m_lo = max(m_lo, seg * len);
    m_hi = min(m_hi, seg * len + len - 1);
  }

  // ta, tb at pixel n of line m in FP64: ta = ka N_u = (a0 + am m) + an n,
  // formed per line from m alone, so that a line's values do not depend on
  // where the walk (or its segment) started.
  const double an = (double)rec.ka * sa, am = (double)rec.ka * ma, a0 = (double)rec.ka * ac0;
  const double bn = (double)rec.kb * sb, bm = (double)rec.kb * mb, b0 = (double)rec.kb * bc0;
  double acc = 0.0;

  // Image reads go through a buffer resource with 32-bit byte offsets (the
  // image is under 4 GiB; reads past its end would return zero).
  const __amdgpu_buffer_rsrc_t img_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(img), (short)0, (int)(4u * (unsigned)g.nx * (unsigned)g.ny), 0x00020000);
  const unsigned bstep = 4u * (unsigned)smin;

  // One line: pixels [flo, fhi] (flo <= fhi) of line m.
  auto walk_line = [&](int m, float flo, float fhi) {
    const int n0 = (int)flo;
    const double dn = (double)flo, dm = (double)m;
    float ta = (float)__builtin_fma(dn, an, __builtin_fma(dm, am, a0));
    float tb = (float)__builtin_fma(dn, bn, __builtin_fma(dm, bm, b0));
    unsigned off = 4u * (unsigned)(m * smaj + n0 * smin);
    const unsigned end = off + (unsigned)((int)fhi - n0 + 1) * bstep;
    float accl = 0.0f;
    do {
      const float x = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(img_rsrc, (int)off, 0, 0));
      accl = __builtin_fmaf(T::conv(rec, ta, tb), x, accl);
      off += bstep;
      ta += dta;
      tb += dtb;
    } while (off != end);
    acc += (double)accl;
  };

  if constexpr (BASIS == BASIS_B1) {
    // Pixel interval of line m: [ceil(lo(m)), floor(hi(m))] from the two
    // pixel-set tests (below), lo and hi affine in m, in FP32.  B1's
    // footprint vanishes (as the cube of the distance) at the edge of the
    // corner range, so a decision flipped by FP32 rounding there changes A
    // by far less than FP32 noise.  sq and sr share a sign whenever the strip
    // is within ~45 degrees of the walk direction, which the per-wave
    // orientation guarantees; otherwise the lane walks whole lines (outside
    // the corner range B1's footprint is zero).
    float lo0 = 0.0f, lo1 = 0.0f, hi0 = (float)(nmin - 1), hi1 = 0.0f;
    if (sq > 0.0 && sr > 0.0) {
      lo0 = (float)((-ra - qc0) * isq);
      lo1 = (float)(-mq * isq);
      hi0 = (float)((rb - rc0) * isr);
      hi1 = (float)(-mr * isr);
    } else if (sq < 0.0 && sr < 0.0) {
      hi0 = (float)((-ra - qc0) * isq);
      hi1 = (float)(-mq * isq);
      lo0 = (float)((rb - rc0) * isr);
      lo1 = (float)(-mr * isr);
    }
    const float lastn = (float)(nmin - 1);
    for (int m = m_lo; m <= m_hi; ++m) {
      const float fm = (float)m;
      const float flo = fmaxf(ceilf(__builtin_fmaf(fm, lo1, lo0)), 0.0f);
      const float fhi = fminf(floorf(__builtin_fmaf(fm, hi1, hi0)), lastn);
      if (flo <= fhi) walk_line(m, flo, fhi);
    }
  } else {
    // Pixel interval of line m, from the two pixel-set tests
    //   q:  N_q + R_q >= 0  <=>  sq n >= -ra - qc0 - mq m
    //   r:  N_r - R_r <  0  <=>  sr n <  rb - rc0 - mr m
    // Each test with a nonzero slope along the walk bounds n from one side
    // (affine in m, evaluated in FP64: B0's model value at the corner-range
    // boundary is not small, so the decision must match an FP64 evaluation
    // of the rule); a test with zero slope bounds the line index instead.
    // Exact integer ties are not special-cased: the corner rule itself is
    // discontinuous there.
    constexpr double kFar = 1e300;
    double lq0 = -kFar, lq1 = 0.0, uq0 = kFar, uq1 = 0.0;
    double lr0 = -kFar, lr1 = 0.0, ur0 = kFar, ur1 = 0.0;
    if (sq > 0.0) {
      lq0 = (-ra - qc0) * isq;
      lq1 = -mq * isq;
    } else if (sq < 0.0) {
      uq0 = (-ra - qc0) * isq;
      uq1 = -mq * isq;
    } else if (mq > 0.0) {
      m_lo = max(m_lo, (int)fmin(ceil((-ra - qc0) / mq), 1e9));
    } else if (mq < 0.0) {
      m_hi = min(m_hi, (int)fmax(floor((-ra - qc0) / mq), -1e9));
    } else if (-ra - qc0 > 0.0) {
      m_hi = -1;
    }
    if (sr > 0.0) {
      ur0 = (rb - rc0) * isr;
      ur1 = -mr * isr;
    } else if (sr < 0.0) {
      lr0 = (rb - rc0) * isr;
      lr1 = -mr * isr;
    } else if (mr > 0.0) {
      m_hi = min(m_hi, (int)fmax(ceil((rb - rc0) / mr) - 1.0, -1e9));
    } else if (mr < 0.0) {
      m_lo = max(m_lo, (int)fmin(floor((rb - rc0) / mr) + 1.0, 1e9));
    } else if (rb - rc0 <= 0.0) {
      m_hi = -1;
    }
    const double lastn = (double)(nmin - 1);
    for (int m = m_lo; m <= m_hi; ++m) {
      const double dm = (double)m;
      const double flo = fmax(fmax(ceil(fma(dm, lq1, lq0)), floor(fma(dm, lr1, lr0)) + 1.0), 0.0);
      const double fhi = fmin(fmin(floor(fma(dm, uq1, uq0)), ceil(fma(dm, ur1, ur0)) - 1.0), lastn);
      if (flo <= fhi) walk_line(m, (float)flo, (float)fhi);
    }
  }
  const double val = acc * g.dx * g.dy * T::scale(rec);
  if (part) part[(size_t)v * g.nu + u] = val;
  else sino[(size_t)v * g.nu + u] = (float)val;
}


// ---------------------------------------------------------------------------
// A^T: pixel-driven gather.
// ---------------------------------------------------------------------------

// Bin window of one view for a pixel tile: bins touched by the tile's box
// widened by h, from the exact projection of its four corners, +-1 margin.
__host__ __device__ inline void tile_window(const Proj2DArgs& g, const ViewGeo& vg, int ix0, int iy0, int* lo,
                                            int* n, int cap = kMaxWindow) {
  double umin = 1e300, umax = -1e300;
  for (int c = 0; c < 4; ++c) {
    double fx = (c & 1) ? (ix0 + kTile - 1 + g.h) : (ix0 - g.h);
    double fy = (c & 2) ? (iy0 + kTile - 1 + g.h) : (iy0 - g.h);
    double wx = fx * g.dx + vg.w0x, wy = fy * g.dy + vg.w0y;
    double n0 = vg.e0x * wy - vg.e0y * wx;
    double be = vg.pux * wy - vg.puy * wx;
    double uu = -n0 / be;
    umin = fmin(umin, uu);
    umax = fmax(umax, uu);
  }
  double l = fmax(floor(umin) - 1.0, 0.0), hgh = fmin(floor(umax) + 1.0, (double)(g.nu - 1));
  int li = (int)l, hi = (int)hgh;
  *lo = li;
  *n = hi >= li ? min(hi - li + 1, cap) : 0;
}

// A^T stages bin records in LDS with the sinogram value (times the footprint
// scale) in place of the scale: stage() writes it, staged_y() reads it.
__device__ __forceinline__ void stage(B1Bin& r, float y) { r.scale *= y; }
__device__ __forceinline__ void stage(B0Bin& r, float y) { r.y = y; }
__device__ __forceinline__ float staged_y(const B1Bin& r) { return r.scale; }
__device__ __forceinline__ float staged_y(const B0Bin& r) { return r.y; }

// USE_LDS: the host has checked that every tile's bin window fits in
// kMaxWindow bins for every view (projector creation); otherwise records are
// read from global memory (L2) directly.
// One 16x16 pixel tile per 256-thread block.  `lds` holds kViewChunk x
// kMaxWindow records (USE_LDS), `win` 2 x kViewChunk ints; every thread of
// the block calls this (it synchronises the block).  Views [v_begin, v_end)
// are gathered; with `part` the tile writes its FP64 sums, already scaled,
// to part[iy * nx + ix], otherwise the finished image values to img.
template <int BASIS, bool USE_LDS>
__device__ __forceinline__ void bwd_pix_tile(const Proj2DArgs& g, const float* __restrict__ sino,
                                             float* __restrict__ img, int tile_x, int tile_y,
                                             typename Traits<BASIS>::Bin (*lds)[kMaxWindow], int* win,
                                             int v_begin = 0, int v_end = -1, double* __restrict__ part = nullptr) {
  if (v_end < 0) v_end = g.np;
  using T = Traits<BASIS>;
  using Bin = typename T::Bin;
  static_assert(sizeof(Bin) % 16 == 0, "bin records are whole float4s");
  int* win_lo = win;
  int* win_n = win + kViewChunk;

  const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x / kTile;
  const int ix0 = tile_x * kTile, iy0 = tile_y * kTile;
  const int ix = ix0 + tx, iy = iy0 + ty;
  const bool valid = ix < g.nx && iy < g.ny;
  const Bin* bins = reinterpret_cast<const Bin*>(g.bins);
  const double fx = ix * g.dx, fy = iy * g.dy;
  const float last = (float)(g.nu - 1);

  double acc = 0.0;
  for (int v0 = v_begin; v0 < v_end; v0 += kViewChunk) {
    const int nv = min(kViewChunk, v_end - v0);
    if (USE_LDS) {
      if ((int)threadIdx.x < nv) {
        int lo, n;
        tile_window(g, g.views[v0 + threadIdx.x], ix0, iy0, &lo, &n);
        win_lo[threadIdx.x] = lo;
        win_n[threadIdx.x] = n;
      }
      __syncthreads();
      for (int idx = threadIdx.x; idx < nv * kMaxWindow; idx += blockDim.x) {
        const int j = idx / kMaxWindow, i = idx - j * kMaxWindow;
        if (i < win_n[j]) {
          const size_t gi = (size_t)(v0 + j) * g.nu + win_lo[j] + i;
          Bin r = bins[gi];
          stage(r, sino[gi]);
          lds[j][i] = r;
        }
      }
      __syncthreads();
    }
    if (valid) {
      for (int j = 0; j < nv; ++j) {
        const int v = v0 + j;
        const ViewGeo& vg = g.views[v];
        const double wx = fx + vg.w0x, wy = fy + vg.w0y;
        const double n0 = vg.e0x * wy - vg.e0y * wx;
        const double be = vg.pux * wy - vg.puy * wx;
        const float bef = (float)be;
        const float kc = -(float)n0 * __builtin_amdgcn_rcpf(bef);
        const float gx = __builtin_fmaf(kc, vg.gx1, vg.gx0), gy = __builtin_fmaf(kc, vg.gy1, vg.gy0);
        const float hw = g.h * (__builtin_fabsf(gx) + __builtin_fabsf(gy)) *
                         __builtin_amdgcn_rcpf(__builtin_fabsf(bef) - vg.db);
        // clamped corner range (layers/A_B1.cu:150-151): pixels projecting
        // wholly before / past the detector still visit bin 0 / nu-1
        float flo = fminf(fmaxf(floorf(kc - hw), 0.0f), last);
        float fhi = fminf(fmaxf(floorf(kc + hw), 0.0f), last);
        if (USE_LDS) {
          // The window holds every bin of the exact corner range of the
          // tile (+-1); bins of the rounded range outside it carry a zero
          // footprint, so clamping keeps the LDS reads inside the window.
          flo = fmaxf(flo, (float)win_lo[j]);
          fhi = fminf(fhi, (float)(win_lo[j] + win_n[j] - 1));
        }
        const int ulo = (int)flo, cnt = (int)fhi - ulo + 1;
        float na = (float)__builtin_fma((double)ulo, be, n0);
        float accv = 0.0f;
        if (USE_LDS) {
          const Bin* rp = &lds[j][ulo - win_lo[j]];
          for (int k = 0; k < cnt; ++k) {
            const Bin rec = rp[k];
            const float nb = na + bef;
            accv = __builtin_fmaf(T::conv(rec, rec.ka * na, rec.kb * nb), staged_y(rec), accv);
            na = nb;
          }
        } else {
          const size_t base = (size_t)v * g.nu + ulo;
          for (int k = 0; k < cnt; ++k) {
            const Bin rec = bins[base + k];
            const float nb = na + bef;
            accv = __builtin_fmaf(T::conv(rec, rec.ka * na, rec.kb * nb), sino[base + k] * T::scale(rec), accv);
            na = nb;
          }
        }
        acc += (double)accv;
      }
    }
    if (USE_LDS) __syncthreads();
  }
  if (valid) {
    const double val = acc * g.dx * g.dy;
    if (part) part[(size_t)iy * g.nx + ix] = val;
    else img[(size_t)iy * g.nx + ix] = (float)val;
  }
}


}  // namespace sgdct
