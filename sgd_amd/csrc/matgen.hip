// matgen.hip -- explicit fan-beam system matrices as COO triplets
// (pixel_matrix_generation / bilinear_matrix_generation,
// layers/A_B0.cu:36-121,449-492 and layers/A_B1.cu:355-427,545-593).
//
// The reference evaluates every (bin, pixel) pair of a view on the GPU into a
// dense nu x (nx ny) FP64 matrix, copies it to the host and scans it.  Here
// one lane owns one (view, bin): it walks the pixels whose footprint can reach
// the bin (the corner rule widened by one pixel, columns ix in ascending
// order, rows iy ascending inside a column) and evaluates the same footprint
// as the projector kernels.  A first pass counts the entries each bin keeps,
// the host turns counts into offsets, a second pass writes
// (p*nu + u, ix*ny + iy, conv) at those offsets -- the reference's order --
// never past the count the first pass reserved.
// An entry is kept when conv > 0 and the pair survives the reference's cull
// "both edge-ray crossings farther than dx in x and dy in y from the pixel
// centre" (A_B0.cu:87-88, A_B1.cu:400-401).
#include <hip/hip_runtime.h>

#include <cmath>

#include "proj2d_dev.hpp"

namespace sgdct {

template <int BASIS, bool WRITE>
__global__ __launch_bounds__(256) void matgen_kernel(Proj2DArgs g, unsigned int* __restrict__ counts,
                                                     const long long* __restrict__ offsets,
                                                     double* __restrict__ out) {
  using T = Traits<BASIS>;
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
  const double acx = -eay * g.dx, acy = eax * g.dy, ac0 = eax * vg.w0y - eay * vg.w0x;
  const double bcx = -eby * g.dx, bcy = ebx * g.dy, bc0 = ebx * vg.w0y - eby * vg.w0x;
  // |x| and |y| components of the unit normal of the bin's central ray: the
  // edge-ray crossings sit at ta * n and tb * n from the pixel centre
  const double cx = eax + 0.5 * vg.pux, cy = eay + 0.5 * vg.puy, cl = sqrt(cx * cx + cy * cy);
  const float nxa = (float)(fabs(cy) / cl), nya = (float)(fabs(cx) / cl);
  // (ta, tb below are in the basis's scaled offset: compare with dx, dy scaled alike)
  const float fdx = (float)g.dx * T::tscale(rec), fdy = (float)g.dy * T::tscale(rec);
  // pixel set: corner rule with half-extent h + 1 (the footprint's support
  // is the projection of the basis at the pixel's depth, within one pixel of
  // the perspective corner range); walk columns ix, rows iy inside
  const double hm = g.h + 1.0;
  const double ra = hm * (fabs(acx) + fabs(acy)), rb = hm * (fabs(bcx) + fabs(bcy));
  unsigned int cnt = 0;
  const long long pos = WRITE ? offsets[(size_t)v * g.nu + u] : 0;
  // the write pass never writes past what the count pass reserved
  const unsigned int cap = WRITE ? counts[(size_t)v * g.nu + u] : 0u;
  for (int ix = 0; ix < g.nx; ++ix) {
    // rows with N_u + ra >= 0 and N_{u+1} - rb < 0 in column ix
    double lo = 0.0, hi = g.ny - 1;
    const double q = acx * ix + ac0 + ra, r = bcx * ix + bc0 - rb;  // + acy iy, + bcy iy
    if (acy > 0.0) lo = fmax(lo, ceil(-q / acy));
    else if (acy < 0.0) hi = fmin(hi, floor(-q / acy));
    else if (q < 0.0) continue;
    if (bcy > 0.0) hi = fmin(hi, ceil(-r / bcy) - 1.0);
    else if (bcy < 0.0) lo = fmax(lo, floor(-r / bcy) + 1.0);
    else if (r >= 0.0) continue;
    lo = fmin(lo, (double)g.ny);
    hi = fmax(hi, -1.0);
    for (int iy = (int)lo; iy <= (int)hi; ++iy) {
      const double na = acx * ix + acy * iy + ac0, nb = bcx * ix + bcy * iy + bc0;
      const float ta = (float)(rec.ka * na), tb = (float)(rec.kb * nb);
      const float fa = __builtin_fabsf(ta), fb = __builtin_fabsf(tb);
      if (fa * nxa > fdx && fb * nxa > fdx && fa * nya > fdy && fb * nya > fdy) continue;
      const float w = T::conv(rec, ta, tb) * T::scale(rec);
      if (!(w > 0.0f)) continue;
      if (WRITE) {
        if (cnt >= cap) break;
        double* o = out + 3 * (pos + cnt);
        o[0] = (double)v * g.nu + u;
        o[1] = (double)ix * g.ny + iy;
        o[2] = (double)w;
      }
      ++cnt;
    }
  }
  if (!WRITE) counts[(size_t)v * g.nu + u] = cnt;
}

hipError_t launch_matgen(int basis, const Proj2DArgs& g, bool write, unsigned int* counts, const long long* offsets,
                         double* out, hipStream_t s) {
  const int waves = g.np * g.nchunk;
  dim3 grid((waves + 3) / 4), block(256);
  if (basis == BASIS_B1) {
    if (write) matgen_kernel<BASIS_B1, true><<<grid, block, 0, s>>>(g, counts, offsets, out);
    else matgen_kernel<BASIS_B1, false><<<grid, block, 0, s>>>(g, counts, offsets, out);
  } else if (basis == BASIS_B0) {
    if (write) matgen_kernel<BASIS_B0, true><<<grid, block, 0, s>>>(g, counts, offsets, out);
    else matgen_kernel<BASIS_B0, false><<<grid, block, 0, s>>>(g, counts, offsets, out);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace sgdct
