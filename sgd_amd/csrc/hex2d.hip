// hex2d.hip -- hexagonal-lattice projectors (H0, Hl), gfx950.
//
// Lattice point (ic, il): row il at y = (il - (nl-1)/2) dyh, column ic at
// x = (ic + (il odd ? 1/2 : 0) - (nc-1)/2) dxh, dxh = sqrt(3) dt; odd rows hold
// nc-1 points (layers/A_H0.cu:172-183, A_Hl.cu:171-182).
//   H0: hexagon indicator (circumradius dt, dyh = 1.5 dt) = three rhombi; the
//       value is the bin average of the sum of the three rhombus projections
//       (box splines of two of va, vb, vc each, A_H0.cu:229-245), divided by 3
//       so that it is the hexagon's projection (the reference omits the 1/3).
//   Hl: three-direction box spline (dyh = dt/2), bin average of
//       box_|va| * box_|vb| * box_|vc| (A_Hl.cu:226-236).
// va, vb, vc = dt sin(theta + 2 pi k / 3), theta the bin's central-ray angle
// atan2(dx, dy) (A_H0.cu:229-232).  Bin ranges come from the six true hexagon
// vertices: the reference's H0 culling uses the Hl row pitch and drops
// off-centre rows (A_H0.cu:82-156 with h0_init's dy = dt/2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "footprint.hpp"
#include "projector.hpp"

namespace sgdct {

hipError_t launch_sum_parts(const double* part, float* out, size_t n, int nparts, hipStream_t s);

// Rhombus (va, vb) of H0: trapezoid box_p * box_q centred at c = (va + vb)/2.
struct H0Bin {
  float ka, kb;
  float c[3], hpq[3], hmq[3], ip[3], i2pq[3];
  float pad[3];
};

// Hl: widths p >= q >= r (p = q + r) of box_p * box_q * box_r.
struct HlBin {
  float ka, kb, p, q, r, inv6pq, inv_r, pad;
};

// Antiderivative of box_p * box_q * box_r (p = q + r), odd in t:
//   S(x) = 1/2 - [(p-x)_+^3 - (q-x)_+^3 - (r-x)_+^3] / (6pqr),   x = |t|,
// written without the cancellations of the difference of cubes:
//   x <= q :  1/2 - (A^2 + AB + B^2)/(6pq) + C^2 (C/r)/(6pq)
//   x >  q :  1/2 - A^2 (A/r)/(6pq)
// with A = (p-x)_+, B = (q-x)_+, C = (r-x)_+  (A - B = r when x <= q).
__device__ __forceinline__ float s_hl(float t, const HlBin& r) {
  const float x = __builtin_fabsf(t);
  const float A = __builtin_fmaxf(r.p - x, 0.0f);
  const float B = __builtin_fmaxf(r.q - x, 0.0f);
  const float C = __builtin_fmaxf(r.r - x, 0.0f);
  const float a2 = A * A;
  const float v1 = (__builtin_fmaf(A, B, a2) + B * B) - C * C * (C * r.inv_r);
  const float v2 = a2 * (A * r.inv_r);
  const float v = ((x <= r.q) ? v1 : v2) * r.inv6pq;
  return copysign_f(0.5f - v, t);
}

template <int HB> struct HexTraits;

template <> struct HexTraits<BASIS_H0> {
  using Bin = H0Bin;
  __device__ static __forceinline__ float conv(const Bin& r, float ta, float tb) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float sa = s_b0(ta - r.c[k], r.hpq[k], r.hmq[k], r.ip[k], r.i2pq[k]);
      const float sb = s_b0(tb - r.c[k], r.hpq[k], r.hmq[k], r.ip[k], r.i2pq[k]);
      s += sb - sa;
    }
    const float cv = s * (1.0f / 3.0f) * __builtin_amdgcn_rcpf(tb - ta);
    return cv > 0.0f ? cv : 0.0f;  // conv > 0 (layers/A_H0.cu:248)
  }
};

template <> struct HexTraits<BASIS_HL> {
  using Bin = HlBin;
  __device__ static __forceinline__ float conv(const Bin& r, float ta, float tb) {
    const float cv = (s_hl(tb, r) - s_hl(ta, r)) * __builtin_amdgcn_rcpf(tb - ta);
    return cv > 0.0f ? cv : 0.0f;  // conv > 0 (layers/A_Hl.cu:238)
  }
};

// max over the six vertices of a pointy-top hexagon of circumradius dt of
// (vertex offset) . (gx, gy)
__device__ __forceinline__ double hex_reach(double gx, double gy, double dt) {
  return dt * fmax(fabs(gy), 0.8660254037844386 * fabs(gx) + 0.5 * fabs(gy));
}

__device__ __forceinline__ float hex_reach_f(float gx, float gy, float dt) {
  return dt * fmaxf(__builtin_fabsf(gy), __builtin_fmaf(0.8660254f, __builtin_fabsf(gx), 0.5f * __builtin_fabsf(gy)));
}

struct HexArgs {
  Proj2DArgs g;  // nx = nc, ny = nl, dx = dxh, dy = dyh
  double dt;
};

// A: bin-driven gather over lattice rows.
// grid (waves / 4, batch, nseg), wave = chunk * np + view.  With nseg > 1,
// segment k = blockIdx.z takes
// lines [k L, k L + L) of its walk (rows or vertical lines, L = ceil(lines /
// nseg)) and writes its FP64
// sum, already scaled, to part[(k * batch + image) * np * nu + v * nu + u];
// sum_parts_kernel adds the segments in order.
// STAGE: the row walk reads the lattice from LDS.  The strips of a wave's 64
// adjacent bins lie along the rows there and each lane walks rows of its own,
// so direct reads touch a cache line per lane and instruction, and the lines
// a wave is working on outgrow the vector L1.  Instead the wave walks the
// lattice in slabs of cw columns (32, halved while the slab's rows do not fit
// in the wave's kHexStage floats): it copies the rows its lanes reach in the
// slab into LDS with coalesced loads, then every lane walks its rows of the
// slab there.  The footprint arithmetic is that of the direct walk; the
// point order within a row is the same, the FP32 partial sums restart at each
// slab.  A slab whose rows do not fit even at the narrowest width is read
// directly.
// Waves whose middle ray is within atan(kHexStageSlope dy / dx) of the rows
// stage (22 degrees on the 0.5 mm Hl lattice); steeper strips cross each row
// in a few points, and the slabs would split their short runs.
constexpr int kHexStage = 2048, kHexCW = 32, kHexCWMin = 8;
constexpr double kHexStageSlope = 1.4;

__device__ __forceinline__ void hex_wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int HB, bool STAGE>
__global__ __launch_bounds__(256) void hex_fwd_kernel(HexArgs ha, const float* __restrict__ img,
                                                      float* __restrict__ sino, double* __restrict__ part) {
  using T = HexTraits<HB>;
  const Proj2DArgs& g = ha.g;
  constexpr int kWaves = 4;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  // chunk-major: the waves of a workgroup take the same 64 bins of adjacent
  // views, so they walk nearly the same lattice points (shared L1 lines) and
  // reach the lattice together or miss it together (a workgroup holds its
  // slots until its longest wave ends; the 74 mm tall Hl lattice leaves three
  // quarters of a view's bins empty near the row direction).  Against
  // view-major order: H0 A 1.094 -> 0.834 ms, Hl 1.258 -> 0.944 ms (296 x 256,
  // 1024 views, 1024 bins, one image); with the LDS staging below Hl 0.651 ms.
  const int chunk = wave / g.np, v = wave - chunk * g.np;
  if (chunk >= g.nchunk) return;
  // with STAGE every lane of the wave takes part in the staging: lanes past
  // the last bin walk nothing and store nothing
  const int u_lane = chunk * 64 + lane;
  const bool lane_ok = u_lane < g.nu;
  if (!STAGE && !lane_ok) return;
  const int u = lane_ok ? u_lane : g.nu - 1;
  const int seg = blockIdx.z, nseg = gridDim.z;
  img += (size_t)blockIdx.y * g.nx * g.ny;
  const ViewGeo vg = g.views[v];
  const typename T::Bin rec = reinterpret_cast<const typename T::Bin*>(g.bins)[(size_t)v * g.nu + u];

  const double eax = vg.e0x + u * vg.pux, eay = vg.e0y + u * vg.puy;
  const double ebx = eax + vg.pux, eby = eay + vg.puy;
  // N(ic', il) = s * ic' + m * il + c, ic' = ic + (il odd) / 2
  const double sa = -eay * g.dx, ma = eax * g.dy, ca = eax * vg.w0y - eay * vg.w0x;
  const double sb = -eby * g.dx, mb = ebx * g.dy, cb = ebx * vg.w0y - eby * vg.w0x;
  // pixel-set edges (clamped end bins see one more bin), as in fwd_bin_kernel
  const double lax = (u == 0) ? eax - vg.pux : eax, lay = (u == 0) ? eay - vg.puy : eay;
  const double lbx = (u == g.nu - 1) ? ebx + vg.pux : ebx, lby = (u == g.nu - 1) ? eby + vg.puy : eby;
  const double sq = -lay * g.dx, mq = lax * g.dy, cq = lax * vg.w0y - lay * vg.w0x;
  const double sr = -lby * g.dx, mr = lbx * g.dy, cr = lbx * vg.w0y - lby * vg.w0x;
  const double rq = hex_reach(-lay, lax, ha.dt), rr = hex_reach(-lby, lbx, ha.dt);

  const double ka = rec.ka, kb = rec.kb;
  double acc = 0.0;
  // Walk lattice rows, or the vertical lines k = 2 ic + (il & 1) (x = k dx / 2,
  // every other row) when the wave's middle ray gives a strip more points per
  // vertical line (|ray_y| dx > 2 dy |ray_x|): steep strips cross every row
  // with only a point or two in each.  The choice is uniform over the wave.
  const int umid = min(chunk * 64 + 32, g.nu - 1);
  const double rmx = vg.e0x + (umid + 0.5) * vg.pux, rmy = vg.e0y + (umid + 0.5) * vg.puy;
  if (fabs(rmy) * g.dx <= 2.0 * g.dy * fabs(rmx)) {
    // rows the strip can reach: between the crossings of its two boundary
    // lines with the first and last column (ic' = 0 and nc), one row of margin
    int il_lo = 0, il_hi = g.ny - 1;
    if (mq != 0.0 && mr != 0.0) {
      const double last = (double)g.nx;
      const double c1 = (-rq - cq) / mq, c2 = (-rq - cq - sq * last) / mq;
      const double c3 = (rr - cr) / mr, c4 = (rr - cr - sr * last) / mr;
      const double lo = fmin(fmin(c1, c2), fmin(c3, c4)), hi = fmax(fmax(c1, c2), fmax(c3, c4));
      il_lo = (int)fmax(floor(lo) - 1.0, 0.0);
      il_hi = (int)fmin(ceil(hi) + 1.0, (double)(g.ny - 1));
    }
    if (nseg > 1) {
      const int len = (g.ny + nseg - 1) / nseg;
      il_lo = max(il_lo, seg * len);
      il_hi = min(il_hi, seg * len + len - 1);
    }
    if (!lane_ok) il_lo = 1, il_hi = 0;
    if (STAGE && fabs(rmy) * g.dx <= kHexStageSlope * g.dy * fabs(rmx)) {
      __shared__ float stage_all[STAGE ? kWaves * kHexStage : 1];
      float* st = stage_all + (threadIdx.x >> 6) * kHexStage;
      const double kas = ka * sa, kbs = kb * sb;
      const double isq = sq != 0.0 ? 1.0 / sq : 0.0, isr = sr != 0.0 ? 1.0 / sr : 0.0;
      const double imq = mq != 0.0 ? 1.0 / mq : 0.0, imr = mr != 0.0 ? 1.0 / mr : 0.0;
      const float dta = (float)kas, dtb = (float)kbs;
      // the lane's rows within the slab ic' in [x0, x1]: between the
      // crossings of its boundary lines with the slab's edges, one row of margin
      auto slab_rows = [&](double x0, double x1, int& lo_r, int& hi_r) {
        lo_r = il_lo;
        hi_r = il_hi;
        if (mq != 0.0 && mr != 0.0) {
          const double c1 = (-rq - cq - sq * x0) * imq, c2 = (-rq - cq - sq * x1) * imq;
          const double c3 = (rr - cr - sr * x0) * imr, c4 = (rr - cr - sr * x1) * imr;
          const double lo = fmin(fmin(c1, c2), fmin(c3, c4)), hi = fmax(fmax(c1, c2), fmax(c3, c4));
          lo_r = max(lo_r, (int)fmin(fmax(floor(lo) - 1.0, -1.0), (double)g.ny));
          hi_r = min(hi_r, (int)fmin(fmax(ceil(hi) + 1.0, -1.0), (double)g.ny));
        }
      };
      for (int c0 = 0; c0 < g.nx;) {
        int cw = kHexCW, lo_r, hi_r, R0, R1;
        for (;;) {
          slab_rows((double)c0, (double)(c0 + cw), lo_r, hi_r);
          const bool any = lo_r <= hi_r;
          R0 = __ockl_wfred_min_i32(any ? lo_r : 0x7fffffff);
          R1 = __ockl_wfred_max_i32(any ? hi_r : -1);
          if (R0 > R1 || (R1 - R0 + 1) * (cw + 1) <= kHexStage || cw == kHexCWMin) break;
          cw >>= 1;
        }
        const int c1 = c0 + cw - 1;
        if (R0 <= R1) {
          const int ls = cw + 1;
          const bool staged = (R1 - R0 + 1) * ls <= kHexStage;  // uniform over the wave
          if (staged) {
            const int lg = __builtin_ctz(cw), n_el = (R1 - R0 + 1) << lg;
            hex_wave_lds_sync();  // the previous slab's reads come first
            for (int e = lane; e < n_el; e += 64) {
              const int r = e >> lg, c = e & (cw - 1);
              st[r * ls + c] = img[(size_t)(R0 + r) * g.nx + min(c0 + c, g.nx - 1)];
            }
            hex_wave_lds_sync();
          }
          for (int il = lo_r; il <= hi_r; ++il) {
            const double sh = (il & 1) ? 0.5 : 0.0;
            const double r1 = -rq - __builtin_fma(mq, (double)il, cq) - sq * sh;  // need sq * ic >= r1
            const double r2 = rr - __builtin_fma(mr, (double)il, cr) - sr * sh;   // need sr * ic <  r2
            double lo = (double)c0, hi = (double)min(c1, g.nx - 1 - (il & 1));
            if (sq > 0.0) lo = fmax(lo, ceil(r1 * isq));
            else if (sq < 0.0) hi = fmin(hi, floor(r1 * isq));
            else if (r1 > 0.0) continue;
            if (sr > 0.0) hi = fmin(hi, ceil(r2 * isr) - 1.0);
            else if (sr < 0.0) lo = fmax(lo, floor(r2 * isr) + 1.0);
            else if (r2 <= 0.0) continue;
            if (lo > hi) continue;
            const int n0 = (int)lo, n1 = (int)hi;
            float ta = (float)__builtin_fma(kas, n0 + sh, ka * __builtin_fma(ma, (double)il, ca));
            float tb = (float)__builtin_fma(kbs, n0 + sh, kb * __builtin_fma(mb, (double)il, cb));
            float accl = 0.0f;
            if (staged) {
              const float* q = st + (il - R0) * ls - c0;
              for (int n = n0; n <= n1; ++n) {
                accl = __builtin_fmaf(T::conv(rec, ta, tb), q[n], accl);
                ta += dta;
                tb += dtb;
              }
            } else {
              const float* p = img + (size_t)il * g.nx;
              for (int n = n0; n <= n1; ++n) {
                accl = __builtin_fmaf(T::conv(rec, ta, tb), p[n], accl);
                ta += dta;
                tb += dtb;
              }
            }
            acc += (double)accl;
          }
        }
        c0 += cw;
      }
    } else {
    // terms linear in the row index, carried from row to row in FP64
    double q_il = mq * il_lo + cq, r_il = mr * il_lo + cr;
    double ta_il = ka * (ma * il_lo + ca), tb_il = kb * (mb * il_lo + cb);
    const double dta_il = ka * ma, dtb_il = kb * mb, kas = ka * sa, kbs = kb * sb;
    const double isq = sq != 0.0 ? 1.0 / sq : 0.0, isr = sr != 0.0 ? 1.0 / sr : 0.0;
    const float dta = (float)kas, dtb = (float)kbs;
    for (int il = il_lo; il <= il_hi; ++il, q_il += mq, r_il += mr, ta_il += dta_il, tb_il += dtb_il) {
      const double sh = (il & 1) ? 0.5 : 0.0;
      const double r1 = -rq - q_il - sq * sh;  // need sq * ic >= r1
      const double r2 = rr - r_il - sr * sh;   // need sr * ic <  r2
      double lo = 0.0, hi = (double)(g.nx - 1 - (il & 1));
      if (sq > 0.0) lo = fmax(lo, ceil(r1 * isq));
      else if (sq < 0.0) hi = fmin(hi, floor(r1 * isq));
      else if (r1 > 0.0) continue;
      if (sr > 0.0) hi = fmin(hi, ceil(r2 * isr) - 1.0);
      else if (sr < 0.0) lo = fmax(lo, floor(r2 * isr) + 1.0);
      else if (r2 <= 0.0) continue;
      if (lo > hi) continue;
      const int n0 = (int)lo, n1 = (int)hi;
      float ta = (float)__builtin_fma(kas, n0 + sh, ta_il);
      float tb = (float)__builtin_fma(kbs, n0 + sh, tb_il);
      const float* p = img + (size_t)il * g.nx;
      float accl = 0.0f;
      // the next point's value is requested one step ahead (the index clamped
      // to the row's last point, so the read needs no branch), which overlaps
      // its latency with the footprint evaluation
      float xn = p[n0];
      for (int n = n0; n <= n1; ++n) {
        const float xc = xn;
        xn = p[min(n + 1, n1)];
        accl = __builtin_fmaf(T::conv(rec, ta, tb), xc, accl);
        ta += dta;
        tb += dtb;
      }
      acc += (double)accl;
    }
    }
  } else {
    // vertical lines the strip can reach: between the crossings of its two
    // boundary lines with the first and last row, one line of margin
    const int kmax = 2 * g.nx - 2;
    int k_lo = 0, k_hi = kmax;
    if (sq != 0.0 && sr != 0.0) {
      const double last = (double)(g.ny - 1);
      const double c1 = (-rq - cq) / sq, c2 = (-rq - cq - mq * last) / sq;
      const double c3 = (rr - cr) / sr, c4 = (rr - cr - mr * last) / sr;
      const double lo = fmin(fmin(c1, c2), fmin(c3, c4)), hi = fmax(fmax(c1, c2), fmax(c3, c4));
      k_lo = (int)fmax(floor(2.0 * lo) - 2.0, 0.0);
      k_hi = (int)fmin(ceil(2.0 * hi) + 2.0, (double)kmax);
    }
    if (nseg > 1) {
      const int len = (kmax + nseg) / nseg;
      k_lo = max(k_lo, seg * len);
      k_hi = min(k_hi, seg * len + len - 1);
    }
    if (!lane_ok) k_lo = 1, k_hi = 0;
    const double imq = mq != 0.0 ? 1.0 / mq : 0.0, imr = mr != 0.0 ? 1.0 / mr : 0.0;
    const double kam = ka * ma, kbm = kb * mb;
    const float dta = (float)(2.0 * kam), dtb = (float)(2.0 * kbm);
    for (int k = k_lo; k <= k_hi; ++k) {
      const int par = k & 1;
      const double xh = 0.5 * k;                   // ic' of the line
      const double r1 = -rq - cq - sq * xh;        // need mq * il >= r1
      const double r2 = rr - cr - sr * xh;         // need mr * il <  r2
      double lo = (double)par, hi = (double)(g.ny - 1);
      if (mq > 0.0) lo = fmax(lo, ceil(r1 * imq));
      else if (mq < 0.0) hi = fmin(hi, floor(r1 * imq));
      else if (r1 > 0.0) continue;
      if (mr > 0.0) hi = fmin(hi, ceil(r2 * imr) - 1.0);
      else if (mr < 0.0) lo = fmax(lo, floor(r2 * imr) + 1.0);
      else if (r2 <= 0.0) continue;
      if (lo > hi) continue;
      int i0 = (int)lo, i1 = (int)hi;
      i0 += (i0 ^ par) & 1;  // rows of the line's parity
      i1 -= (i1 ^ par) & 1;
      if (i0 > i1) continue;
      float ta = (float)__builtin_fma(kam, (double)i0, ka * (sa * xh + ca));
      float tb = (float)__builtin_fma(kbm, (double)i0, kb * (sb * xh + cb));
      const float* p = img + (k >> 1);
      float accl = 0.0f;
      float xn = p[(size_t)i0 * g.nx];
      for (int il = i0; il <= i1; il += 2) {
        const float xc = xn;
        xn = p[(size_t)min(il + 2, i1) * g.nx];
        accl = __builtin_fmaf(T::conv(rec, ta, tb), xc, accl);
        ta += dta;
        tb += dtb;
      }
      acc += (double)accl;
    }
  }
  if (!lane_ok) return;
  const size_t out = (size_t)v * g.nu + u, nsino = (size_t)g.np * g.nu;
  if (nseg > 1) part[((size_t)seg * gridDim.y + blockIdx.y) * nsino + out] = acc * g.dx * g.dy;
  else sino[blockIdx.y * nsino + out] = (float)(acc * g.dx * g.dy);
}

constexpr int kHexH0Window = 72;

// A^T: lattice-point-driven gather, 16 x 16 point tiles, LDS bin windows.
// grid (tiles, tiles, batch * ngroup), H0's blocks mapped by xcd_block: z =
// b * ngroup + k takes views
// [k vpg, (k+1) vpg); with ngroup > 1 it writes its FP64 sum to
// part[(k * batch + b) * nc * nl + point], added in group order afterwards.
template <int HB>
__global__ __launch_bounds__(256) void hex_bwd_kernel(HexArgs ha, const float* __restrict__ sino,
                                                      float* __restrict__ img, int ngroup, int vpg,
                                                      double* __restrict__ part) {
  using T = HexTraits<HB>;
  using Bin = typename T::Bin;
  constexpr int BIN4 = sizeof(Bin) / 16;
  static_assert(sizeof(Bin) % 16 == 0, "bin records are whole float4s");
  // LDS bin window per view: H0's 80-byte records at 128 bins would hold 42
  // KB per block (3 waves per SIMD); a 16 x 16 H0 tile spans up to ~66 bins
  // at 0.3 mm per bin, and bins past the window are read from global memory
  constexpr int kWin = HB == BASIS_H0 ? kHexH0Window : kMaxWindow;
  __shared__ float4 lds[kViewChunk][kWin][BIN4];
  __shared__ float ldsy[kViewChunk][kWin];
  __shared__ int win_lo[kViewChunk], win_n[kViewChunk];
  const Proj2DArgs& g = ha.g;
  // XCD-contiguous blocks (xcd_block) for H0: A^T 0.959 -> 0.939 ms (296 x
  // 256, 1024 views, 1024 bins); Hl was slower with them (0.547 -> 0.562 ms)
  const uint3 blk = HB == BASIS_H0 ? xcd_block() : make_uint3(blockIdx.x, blockIdx.y, blockIdx.z);
  const int ic0 = blk.x * kTile, il0 = blk.y * kTile;
  const int ic = ic0 + (threadIdx.x & (kTile - 1)), il = il0 + threadIdx.x / kTile;
  const int k = blk.z % ngroup, b = blk.z / ngroup, batch = gridDim.z / ngroup;
  const size_t npt = (size_t)g.nx * g.ny;
  sino += (size_t)b * g.np * g.nu;
  img += b * npt;
  const int v_begin = k * vpg, v_end = min(g.np, v_begin + vpg);
  const bool inside = ic < g.nx && il < g.ny;
  const bool valid = inside && !((il & 1) && ic == g.nx - 1);
  const Bin* bins = reinterpret_cast<const Bin*>(g.bins);
  const double sh = (il & 1) ? 0.5 : 0.0;

  double acc = 0.0;
  for (int v0 = v_begin; v0 < v_end; v0 += kViewChunk) {
    const int nv = min(kViewChunk, v_end - v0);
    if ((int)threadIdx.x < nv) {
      // window: the tile's bounding box widened by one circumradius
      const ViewGeo& vg = g.views[v0 + threadIdx.x];
      double umin = 1e300, umax = -1e300;
      for (int c = 0; c < 4; ++c) {
        double fx = (c & 1) ? (ic0 + kTile + 0.5) * g.dx + ha.dt : (ic0 - 0.5) * g.dx - ha.dt;
        double fy = (c & 2) ? (il0 + kTile) * g.dy + ha.dt : (il0 - 1) * g.dy - ha.dt;
        double wx = fx + vg.w0x, wy = fy + vg.w0y;
        double uu = -(vg.e0x * wy - vg.e0y * wx) / (vg.pux * wy - vg.puy * wx);
        umin = fmin(umin, uu);
        umax = fmax(umax, uu);
      }
      int lo = (int)fmax(floor(umin) - 1.0, 0.0), hi = (int)fmin(floor(umax) + 1.0, (double)(g.nu - 1));
      win_lo[threadIdx.x] = lo;
      win_n[threadIdx.x] = hi >= lo ? min(hi - lo + 1, kWin) : 0;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nv * kWin; idx += blockDim.x) {
      const int j = idx / kWin, i = idx - j * kWin;
      if (i < win_n[j]) {
        const size_t gi = (size_t)(v0 + j) * g.nu + win_lo[j] + i;
        const float4* src = reinterpret_cast<const float4*>(bins + gi);
#pragma unroll
        for (int k = 0; k < BIN4; ++k) lds[j][i][k] = src[k];
        ldsy[j][i] = sino[gi];
      }
    }
    __syncthreads();
    if (valid) {
      for (int j = 0; j < nv; ++j) {
        const int v = v0 + j;
        const ViewGeo& vg = g.views[v];
        const double wx = (ic + sh) * g.dx + vg.w0x, wy = il * g.dy + vg.w0y;
        const double n0 = vg.e0x * wy - vg.e0y * wx;
        const double be = vg.pux * wy - vg.puy * wx;
        const float bef = (float)be;
        // bin range in FP32 (it only bounds the bins visited; a bin gained or
        // lost at a rounding boundary carries a vanishing footprint)
        const float kc = -(float)n0 * __builtin_amdgcn_rcpf(bef);
        // k*(vertex) - kc = -(d . G) / (beta + d . grad beta), G = grad N_0 + kc grad beta
        const float gx = -__builtin_fmaf(kc, (float)vg.puy, (float)vg.e0y);
        const float gy = __builtin_fmaf(kc, (float)vg.pux, (float)vg.e0x);
        const float dt = (float)ha.dt;
        const float hw = hex_reach_f(gx, gy, dt) *
                         __builtin_amdgcn_rcpf(__builtin_fabsf(bef) - hex_reach_f(-(float)vg.puy, (float)vg.pux, dt));
        const float last = (float)(g.nu - 1);
        const int ulo = (int)fminf(fmaxf(floorf(kc - hw), 0.0f), last);
        const int uhi = (int)fminf(fmaxf(floorf(kc + hw), 0.0f), last);
        float na = (float)(n0 + (double)ulo * be);
        const int wlo = win_lo[j], wn = win_n[j];
        float accv = 0.0f;
        for (int u = ulo; u <= uhi; ++u) {
          const int wi = u - wlo;
          Bin rec;
          float yv;
          if (wi >= 0 && wi < wn) {
            float4* d = reinterpret_cast<float4*>(&rec);
#pragma unroll
            for (int k = 0; k < BIN4; ++k) d[k] = lds[j][wi][k];
            yv = ldsy[j][wi];
          } else {
            rec = bins[(size_t)v * g.nu + u];
            yv = sino[(size_t)v * g.nu + u];
          }
          const float nb = na + bef;
          accv = __builtin_fmaf(T::conv(rec, rec.ka * na, rec.kb * nb), yv, accv);
          na = nb;
        }
        acc += (double)accv;
      }
    }
    __syncthreads();
  }
  if (inside) {
    const double val = valid ? acc * g.dx * g.dy : 0.0;
    if (ngroup > 1) part[(size_t)k * batch * npt + b * npt + (size_t)il * g.nx + ic] = val;
    else img[(size_t)il * g.nx + ic] = (float)val;
  }
}

// Evaluation count (clamped six-vertex ranges).
__global__ __launch_bounds__(256) void hex_count_kernel(HexArgs ha, unsigned long long* out) {
  __shared__ unsigned long long part[256];
  const Proj2DArgs& g = ha.g;
  const int ic = blockIdx.x * 16 + (threadIdx.x & 15), il = blockIdx.y * 16 + (threadIdx.x >> 4);
  unsigned long long c = 0;
  if (ic < g.nx && il < g.ny && !((il & 1) && ic == g.nx - 1)) {
    const double cx = (ic + ((il & 1) ? 0.5 : 0.0)) * g.dx, cy = il * g.dy;
    for (int v = 0; v < g.np; ++v) {
      const ViewGeo& vg = g.views[v];
      double umin = 1e300, umax = -1e300;
      for (int k = 0; k < 6; ++k) {
        const double a = 1.5707963267948966 + k * 1.0471975511965976;
        const double wx = cx + ha.dt * cos(a) + vg.w0x, wy = cy + ha.dt * sin(a) + vg.w0y;
        const double uu = -(vg.e0x * wy - vg.e0y * wx) / (vg.pux * wy - vg.puy * wx);
        umin = fmin(umin, uu);
        umax = fmax(umax, uu);
      }
      const int lo = (int)fmin(fmax(floor(umin), 0.0), (double)(g.nu - 1));
      const int hi = (int)fmin(fmax(floor(umax), 0.0), (double)(g.nu - 1));
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

// Per-(view, bin) constants (host side, FP64).
void hex_bin_constants(int basis, const ViewGeo& v, int nu, double dt, void* out) {
  for (int u = 0; u < nu; ++u) {
    const double ax = v.e0x + u * v.pux, ay = v.e0y + u * v.puy;
    const double bx = ax + v.pux, by = ay + v.puy;
    const double cx = ax + 0.5 * v.pux, cy = ay + 0.5 * v.puy;
    const double len = std::sqrt(cx * cx + cy * cy);
    const double ka = -len / (ax * cx + ay * cy), kb = -len / (bx * cx + by * cy);
    // central-ray angle, rounded to FP32 like atan2f (layers/A_H0.cu:229)
    const double th = (double)(float)std::atan2(cx, cy);
    double vk[3];
    for (int k = 0; k < 3; ++k) vk[k] = dt * std::sin(th + k * 2.0 * 3.14159265358979323846 / 3.0);
    if (basis == BASIS_H0) {
      H0Bin r{};
      r.ka = (float)ka;
      r.kb = (float)kb;
      const int pr[3][2] = {{0, 1}, {1, 2}, {2, 0}};
      for (int k = 0; k < 3; ++k) {
        const double a = vk[pr[k][0]], b = vk[pr[k][1]];
        const double p = std::max(std::fabs(a), std::fabs(b));
        double q = std::min(std::fabs(a), std::fabs(b));
        if (!(q > 1e-7)) q = 0.0;
        r.c[k] = (float)(0.5 * (a + b));
        r.hpq[k] = (float)(0.5 * (p + q));
        r.hmq[k] = (float)(0.5 * (p - q));
        r.ip[k] = (float)(1.0 / p);
        r.i2pq[k] = q > 0.0 ? (float)(1.0 / (2.0 * p * q)) : 0.0f;
      }
      static_cast<H0Bin*>(out)[u] = r;
    } else {
      HlBin r{};
      r.ka = (float)ka;
      r.kb = (float)kb;
      double w[3] = {std::fabs(vk[0]), std::fabs(vk[1]), std::fabs(vk[2])};
      std::sort(w, w + 3);
      const double p = w[2], q = w[1], rr = w[0];
      r.p = (float)p;
      r.q = (float)q;
      r.r = (float)rr;
      r.inv6pq = (float)(1.0 / (6.0 * p * q));
      r.inv_r = rr > 1e-30 ? (float)(1.0 / rr) : 0.0f;
      static_cast<HlBin*>(out)[u] = r;
    }
  }
}

size_t hex_bin_size(int basis) { return basis == BASIS_H0 ? sizeof(H0Bin) : sizeof(HlBin); }

// nseg > 1 needs part with nseg * batch * np * nu doubles.
hipError_t launch_hex_forward(int basis, const Proj2DArgs& g, double dt, const float* img, float* sino, int batch,
                              int nseg, double* part, hipStream_t s) {
  if (nseg < 1 || (nseg > 1 && !part)) return hipErrorInvalidValue;
  HexArgs ha{g, dt};
  const int waves = g.np * g.nchunk;
  if (basis == BASIS_H0)
    hex_fwd_kernel<BASIS_H0, false><<<dim3((waves + 3) / 4, batch, nseg), dim3(256), 0, s>>>(ha, img, sino, part);
  else
    hex_fwd_kernel<BASIS_HL, true><<<dim3((waves + 3) / 4, batch, nseg), dim3(256), 0, s>>>(ha, img, sino, part);
  if (nseg > 1) return launch_sum_parts(part, sino, (size_t)batch * g.np * g.nu, nseg, s);
  return hipGetLastError();
}

hipError_t launch_hex_backward(int basis, const Proj2DArgs& g, double dt, const float* sino, float* img, int batch,
                               int ngroup, double* part, hipStream_t s) {
  if (ngroup < 1 || (ngroup > 1 && !part)) return hipErrorInvalidValue;
  HexArgs ha{g, dt};
  const int vpg = (g.np + ngroup - 1) / ngroup;
  dim3 grid((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile, batch * ngroup), block(256);
  if (basis == BASIS_H0) hex_bwd_kernel<BASIS_H0><<<grid, block, 0, s>>>(ha, sino, img, ngroup, vpg, part);
  else hex_bwd_kernel<BASIS_HL><<<grid, block, 0, s>>>(ha, sino, img, ngroup, vpg, part);
  if (ngroup > 1) return launch_sum_parts(part, img, (size_t)batch * g.nx * g.ny, ngroup, s);
  return hipGetLastError();
}

// Explicit system matrix (h0_matrix_generation / hl_matrix_generation; the
// reference's hl_matrix_generate, layers/A_Hl.cu:370-451, and its H0 twin,
// A_H0.cu:360, are never called): one lane per (view, bin) walks every
// lattice row the bin's strip can reach, points in ascending column order,
// with the same pixel-set tests and footprint as hex_fwd_kernel, and keeps
// the points whose footprint is > 0 as (p nu + u, il nc + ic, conv) --
// without the dx dy factor, as the reference's generators.  A count pass,
// host offsets, then a write pass that never writes past its count.
template <int HB, bool WRITE>
__global__ __launch_bounds__(256) void hex_matgen_kernel(HexArgs ha, unsigned int* __restrict__ counts,
                                                         const long long* __restrict__ offsets,
                                                         double* __restrict__ out) {
  using T = HexTraits<HB>;
  const Proj2DArgs& g = ha.g;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int v = wave / g.nchunk;
  if (v >= g.np) return;
  const int u = (wave - v * g.nchunk) * 64 + lane;
  if (u >= g.nu) return;
  const ViewGeo vg = g.views[v];
  const typename T::Bin rec = reinterpret_cast<const typename T::Bin*>(g.bins)[(size_t)v * g.nu + u];
  const double eax = vg.e0x + u * vg.pux, eay = vg.e0y + u * vg.puy;
  const double ebx = eax + vg.pux, eby = eay + vg.puy;
  const double sa = -eay * g.dx, ma = eax * g.dy, ca = eax * vg.w0y - eay * vg.w0x;
  const double sb = -eby * g.dx, mb = ebx * g.dy, cb = ebx * vg.w0y - eby * vg.w0x;
  const double lax = (u == 0) ? eax - vg.pux : eax, lay = (u == 0) ? eay - vg.puy : eay;
  const double lbx = (u == g.nu - 1) ? ebx + vg.pux : ebx, lby = (u == g.nu - 1) ? eby + vg.puy : eby;
  const double sq = -lay * g.dx, mq = lax * g.dy, cq = lax * vg.w0y - lay * vg.w0x;
  const double sr = -lby * g.dx, mr = lbx * g.dy, cr = lbx * vg.w0y - lby * vg.w0x;
  const double rq = hex_reach(-lay, lax, ha.dt), rr = hex_reach(-lby, lbx, ha.dt);
  const double ka = rec.ka, kb = rec.kb;
  const double isq = sq != 0.0 ? 1.0 / sq : 0.0, isr = sr != 0.0 ? 1.0 / sr : 0.0;
  const size_t bin = (size_t)v * g.nu + u;
  const long long pos = WRITE ? offsets[bin] : 0;
  const unsigned int cap = WRITE ? counts[bin] : 0u;
  unsigned int cnt = 0;
  for (int il = 0; il < g.ny; ++il) {
    const double sh = (il & 1) ? 0.5 : 0.0;
    const double r1 = -rq - (mq * il + cq) - sq * sh;
    const double r2 = rr - (mr * il + cr) - sr * sh;
    double lo = 0.0, hi = (double)(g.nx - 1 - (il & 1));
    if (sq > 0.0) lo = fmax(lo, ceil(r1 * isq));
    else if (sq < 0.0) hi = fmin(hi, floor(r1 * isq));
    else if (r1 > 0.0) continue;
    if (sr > 0.0) hi = fmin(hi, ceil(r2 * isr) - 1.0);
    else if (sr < 0.0) lo = fmax(lo, floor(r2 * isr) + 1.0);
    else if (r2 <= 0.0) continue;
    for (int ic = (int)lo; ic <= (int)hi; ++ic) {
      const double xc = ic + sh;
      const float ta = (float)(ka * (sa * xc + ma * il + ca));
      const float tb = (float)(kb * (sb * xc + mb * il + cb));
      const float w = T::conv(rec, ta, tb);
      if (!(w > 0.0f)) continue;
      if (WRITE) {
        if (cnt >= cap) return;
        double* o = out + 3 * (pos + cnt);
        o[0] = (double)bin;
        o[1] = (double)il * g.nx + ic;
        o[2] = (double)w;
      }
      ++cnt;
    }
  }
  if (!WRITE) counts[bin] = cnt;
}

hipError_t launch_hex_matgen(int basis, const Proj2DArgs& g, double dt, bool write, unsigned int* counts,
                             const long long* offsets, double* out, hipStream_t s) {
  HexArgs ha{g, dt};
  dim3 grid((g.np * g.nchunk + 3) / 4), block(256);
  if (basis == BASIS_H0) {
    if (write) hex_matgen_kernel<BASIS_H0, true><<<grid, block, 0, s>>>(ha, counts, offsets, out);
    else hex_matgen_kernel<BASIS_H0, false><<<grid, block, 0, s>>>(ha, counts, offsets, out);
  } else if (basis == BASIS_HL) {
    if (write) hex_matgen_kernel<BASIS_HL, true><<<grid, block, 0, s>>>(ha, counts, offsets, out);
    else hex_matgen_kernel<BASIS_HL, false><<<grid, block, 0, s>>>(ha, counts, offsets, out);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_hex_count(const Proj2DArgs& g, double dt, unsigned long long* out, hipStream_t s) {
  HexArgs ha{g, dt};
  hex_count_kernel<<<dim3((g.nx + 15) / 16, (g.ny + 15) / 16), dim3(256), 0, s>>>(ha, out);
  return hipGetLastError();
}

}  // namespace sgdct
