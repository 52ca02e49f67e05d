// capi.cpp -- C ABI (include/sgdct.h): handles, host-side FP64 geometry
// tables, host-array entry points with the reference's layouts, and FP32
// device-pointer entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sgdct.h"
#include "footprint.hpp"
#include "projector.hpp"

namespace sgdct {
hipError_t launch_forward(int basis, const Proj2DArgs& g, const float* img, float* sino, int batch,
                          hipStream_t s);
hipError_t launch_backward(int basis, const Proj2DArgs& g, bool use_lds, const float* sino, float* img, int batch,
                           hipStream_t s);
int max_tile_window(const Proj2DArgs& g, const ViewGeo* views_host);
hipError_t launch_count(const Proj2DArgs& g, unsigned long long* out, hipStream_t s);
hipError_t launch_sf_forward(const Proj2DArgs& g, const float* img, float* sino, int batch, hipStream_t s);
hipError_t launch_sf_backward(const Proj2DArgs& g, const float* sino, float* img, int batch, hipStream_t s);
void hex_bin_constants(int basis, const ViewGeo& v, int nu, double dt, void* out);
size_t hex_bin_size(int basis);
hipError_t launch_hex_forward(int basis, const Proj2DArgs& g, double dt, const float* img, float* sino, int batch,
                              hipStream_t s);
hipError_t launch_hex_backward(int basis, const Proj2DArgs& g, double dt, const float* sino, float* img, int batch,
                               hipStream_t s);
hipError_t launch_hex_count(const Proj2DArgs& g, double dt, unsigned long long* out, hipStream_t s);
hipError_t launch_tv_grad(const float* x, float* g, int nx, int ny, int batch, float eps, hipStream_t s);
hipError_t launch_resid(const float* b, const float* ax, const float* w, float* r, size_t n, size_t per, hipStream_t s);
hipError_t launch_update(float* x, const float* c, const float* v, float lam, int positivity, size_t n, size_t per,
                         hipStream_t s);
hipError_t launch_axpy(float* x, const float* g, float a, size_t n, hipStream_t s);
}  // namespace sgdct

using namespace sgdct;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? SGDCT_ENOMEM : SGDCT_EDEVICE,
              std::string(what) + ": " + hipGetErrorString(e));
}

// Source, detector centre and detector step of one view, utils/geo.cu:240-262:
// rotation by the view angle, with angles within 1e-7 of the 45-degree
// diagonals snapped onto them (utils/geo.cu:254-257).
void view_placement(double angle, double lsd, double lso, double du, double drho, double src[2],
                    double dtv[2], double pu[2]) {
  const double kPi = 3.14159265358979323846;
  double beta = angle, s = std::sin(beta), c = std::cos(beta);
  if (std::fabs(s - c) < 1e-7) beta = (s > 0) ? kPi / 4 : -kPi * 3 / 4;
  else if (std::fabs(s + c) < 1e-7) beta = (s > 0) ? kPi * 3 / 4 : -kPi / 4;
  s = std::sin(beta);
  c = std::cos(beta);
  src[0] = (-lso + drho) * c;
  src[1] = (-lso + drho) * s;
  dtv[0] = (lsd - lso) * c;
  dtv[1] = (lsd - lso) * s;
  pu[0] = -du * s;
  pu[1] = du * c;
}

}  // namespace

struct sgdct_projector {
  int basis = 0;
  int nx = 0, ny = 0, np = 0, nu = 0;
  double dx = 0, dy = 0, du = 0, lsd = 0, lso = 0;
  double dt = 0;  // hexagonal lattices: circumradius / lattice constant
  float h = 0.5f;
  int device = 0;
  ViewGeo* d_views = nullptr;
  void* d_bins = nullptr;
  float* d_img = nullptr;   // staging for the host-array entry points
  float* d_sino = nullptr;
  float* h_img = nullptr;   // pinned
  float* h_sino = nullptr;  // pinned
  unsigned long long* d_count = nullptr;
  long long evals = -1;
  bool bwd_lds = false;  // A^T stages bin windows in LDS (B0 / B1)
  Proj2DArgs args{};

  ~sgdct_projector() {
    if (d_views) (void)hipFree(d_views);
    if (d_bins) (void)hipFree(d_bins);
    if (d_img) (void)hipFree(d_img);
    if (d_sino) (void)hipFree(d_sino);
    if (d_count) (void)hipFree(d_count);
    if (h_img) (void)hipHostFree(h_img);
    if (h_sino) (void)hipHostFree(h_sino);
  }
};

namespace {

// Per-(view, bin) constants, FP64 on the host, stored FP32 like the
// reference's per-bin vectors (layers/A_B1.cu:429-478, A_B0.cu:338-388).
void bin_constants(int basis, const ViewGeo& v, int nu, double dx, double dy, void* out) {
  for (int u = 0; u < nu; ++u) {
    const double ax = v.e0x + u * v.pux, ay = v.e0y + u * v.puy;        // E_u - p
    const double bx = ax + v.pux, by = ay + v.puy;                      // E_{u+1} - p
    const double cx = ax + 0.5 * v.pux, cy = ay + 0.5 * v.puy;          // s_u - p
    const double len = std::sqrt(cx * cx + cy * cy);
    const double ka = -len / (ax * cx + ay * cy), kb = -len / (bx * cx + by * cy);
    const double th = std::atan2(std::fabs(cx), std::fabs(cy));
    const float vx = (float)std::fabs(dx * std::cos(th)), vy = (float)std::fabs(dy * std::sin(th));
    if (basis == BASIS_B1) {
      B1Bin r{};
      r.ka = (float)ka;
      r.kb = (float)kb;
      const double a = std::max(vx, vy);
      double b = std::min(vx, vy);
      if (b > 1e-7) {
        const double c = 1.0 / (24.0 * a * a * b * b);
        r.k2 = (float)(1.0 / (2.0 * a * a));
        r.k4 = (float)(b * b / (12.0 * a * a));
        r.cq = (float)c;
        r.c2 = (float)(2.0 * c);
      } else {  // 1tri1box limit, spline_conv.hh:393-405
        b = 0.0;
        r.k2 = (float)(1.0 / (2.0 * a * a));
        r.k4 = 0.0f;
        r.cq = 0.0f;
        r.c2 = 0.0f;
      }
      r.a = (float)a;
      r.b = (float)b;
      r.apb = (float)(a + b);
      static_cast<B1Bin*>(out)[u] = r;
    } else {
      B0Bin r{};
      r.ka = (float)ka;
      r.kb = (float)kb;
      const double p = std::max(vx, vy);
      double q = std::min(vx, vy);
      if (!(q > 1e-7)) q = 0.0;  // two-direction limit, spline_conv.hh:239-242
      r.hpq = (float)(0.5 * (p + q));
      r.hmq = (float)(0.5 * (p - q));
      r.inv_p = (float)(1.0 / p);
      r.inv_2pq = q > 0.0 ? (float)(1.0 / (2.0 * p * q)) : 0.0f;
      static_cast<B0Bin*>(out)[u] = r;
    }
  }
}

bool is_hex(int basis) { return basis == BASIS_H0 || basis == BASIS_HL; }

sgdct_projector* create(int basis, int nx, int ny, int np, int nu, double dx, double dy, double du,
                        double lsd, double lso, const double* angles, const double* drho, double dt = 0.0) {
  if (nx <= 0 || ny <= 0 || np <= 0 || nu <= 0 || !(dx > 0) || !(dy > 0) || !(du > 0) ||
      !(lsd > 0) || !(lso > 0) || !angles) {
    fail(SGDCT_EINVAL, "invalid geometry arguments");
    return nullptr;
  }
  auto* h = new (std::nothrow) sgdct_projector();
  if (!h) {
    fail(SGDCT_ENOMEM, "host allocation failed");
    return nullptr;
  }
  h->basis = basis;
  h->nx = nx; h->ny = ny; h->np = np; h->nu = nu;
  h->dx = dx; h->dy = dy; h->du = du; h->lsd = lsd; h->lso = lso;
  h->h = (basis == BASIS_B1) ? 1.0f : 0.5f;
  h->dt = dt;

  std::vector<ViewGeo> views(np);
  const double cxo = 0.5 * (nx - 1) * dx, cyo = 0.5 * (ny - 1) * dy;
  for (int p = 0; p < np; ++p) {
    double src[2], dtv[2], pu[2];
    view_placement(angles[p], lsd, lso, du, drho ? drho[p] : 0.0, src, dtv, pu);
    ViewGeo& v = views[p];
    v.px = src[0];
    v.py = src[1];
    v.pux = pu[0];
    v.puy = pu[1];
    v.e0x = dtv[0] - 0.5 * nu * pu[0] - src[0];
    v.e0y = dtv[1] - 0.5 * nu * pu[1] - src[1];
    v.w0x = -cxo - src[0];
    v.w0y = -cyo - src[1];
    v.gx0 = (float)(-v.e0y * dx);
    v.gy0 = (float)(v.e0x * dy);
    v.gx1 = (float)(-v.puy * dx);
    v.gy1 = (float)(v.pux * dy);
    v.db = (float)(h->h * (std::fabs(v.puy * dx) + std::fabs(v.pux * dy)));
  }
  size_t rec = 0;
  if (basis == BASIS_B1) rec = sizeof(B1Bin);
  else if (basis == BASIS_B0) rec = sizeof(B0Bin);
  else if (is_hex(basis)) rec = hex_bin_size(basis);
  std::vector<unsigned char> bins(rec * (size_t)np * nu);
  for (int p = 0; p < np && rec; ++p) {
    unsigned char* dst = bins.data() + rec * (size_t)p * nu;
    if (is_hex(basis)) hex_bin_constants(basis, views[p], nu, dt, dst);
    else bin_constants(basis, views[p], nu, dx, dy, dst);
  }

  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipMalloc(&h->d_views, sizeof(ViewGeo) * np);
  if (e == hipSuccess) e = hipMemcpy(h->d_views, views.data(), sizeof(ViewGeo) * np, hipMemcpyHostToDevice);
  if (e == hipSuccess && rec) e = hipMalloc(&h->d_bins, bins.size());
  if (e == hipSuccess && rec) e = hipMemcpy(h->d_bins, bins.data(), bins.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hip_fail(e, "projector init");
    delete h;
    return nullptr;
  }
  Proj2DArgs& a = h->args;
  a.views = h->d_views;
  a.bins = h->d_bins;
  a.nx = nx; a.ny = ny; a.np = np; a.nu = nu;
  a.dx = dx; a.dy = dy;
  a.h = h->h;
  a.nchunk = (nu + 63) / 64;
  // 2 bins of slack for FP64 contraction differences between host and device
  if (basis == BASIS_B0 || basis == BASIS_B1) h->bwd_lds = max_tile_window(a, views.data()) <= kMaxWindow - 2;
  return h;
}

int ensure_staging(sgdct_projector* h) {
  if (h->d_img) return SGDCT_OK;
  const size_t ni = (size_t)h->nx * h->ny, ns = (size_t)h->np * h->nu;
  hipError_t e = hipMalloc(&h->d_img, ni * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&h->d_sino, ns * sizeof(float));
  if (e == hipSuccess) e = hipHostMalloc(&h->h_img, ni * sizeof(float));
  if (e == hipSuccess) e = hipHostMalloc(&h->h_sino, ns * sizeof(float));
  if (e != hipSuccess) return hip_fail(e, "staging allocation");
  return SGDCT_OK;
}

int run_forward(sgdct_projector* h, const float* img, float* sino, int batch, hipStream_t s) {
  hipError_t e;
  if (is_hex(h->basis)) e = launch_hex_forward(h->basis, h->args, h->dt, img, sino, batch, s);
  else if (h->basis == BASIS_SF) e = launch_sf_forward(h->args, img, sino, batch, s);
  else e = launch_forward(h->basis, h->args, img, sino, batch, s);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "forward launch");
}

int run_backward(sgdct_projector* h, const float* sino, float* img, int batch, hipStream_t s) {
  hipError_t e;
  if (is_hex(h->basis)) e = launch_hex_backward(h->basis, h->args, h->dt, sino, img, batch, s);
  else if (h->basis == BASIS_SF) e = launch_sf_backward(h->args, sino, img, batch, s);
  else e = launch_backward(h->basis, h->args, h->bwd_lds, sino, img, batch, s);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "backward launch");
}

long long eval_count(sgdct_projector* h) {
  if (h->evals >= 0) return h->evals;
  hipError_t e = hipSuccess;
  if (!h->d_count) e = hipMalloc(&h->d_count, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(h->d_count, 0, sizeof(unsigned long long));
  if (e == hipSuccess)
    e = is_hex(h->basis) ? launch_hex_count(h->args, h->dt, h->d_count, nullptr) : launch_count(h->args, h->d_count, nullptr);
  unsigned long long c = 0;
  if (e == hipSuccess) e = hipMemcpy(&c, h->d_count, sizeof(c), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "evaluation count");
  h->evals = (long long)c;
  return h->evals;
}

// Host-array forward: x[ix*ny+iy] (FP64) -> b[p*nu+u] (FP64).
int host_forward(double* b, double* x, sgdct_projector* h) {
  if (!h || !b || !x) return fail(SGDCT_EINVAL, "null argument");
  try {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "set device");
    int rc = ensure_staging(h);
    if (rc) return rc;
    const int nx = h->nx, ny = h->ny;
    for (int ix = 0; ix < nx; ++ix)
      for (int iy = 0; iy < ny; ++iy) h->h_img[(size_t)iy * nx + ix] = (float)x[(size_t)ix * ny + iy];
    const size_t ni = (size_t)nx * ny, ns = (size_t)h->np * h->nu;
    e = hipMemcpy(h->d_img, h->h_img, ni * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "upload image");
    rc = run_forward(h, h->d_img, h->d_sino, 1, nullptr);
    if (rc) return rc;
    e = hipMemcpy(h->h_sino, h->d_sino, ns * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "download sinogram");
    for (size_t i = 0; i < ns; ++i) b[i] = (double)h->h_sino[i];
    long long c = eval_count(h);
    if (c < 0) return (int)c;
    return (int)std::min<long long>(c, INT_MAX);
  } catch (...) {
    return fail(SGDCT_EDEVICE, "unexpected exception in forward");
  }
}

// Host-array back-projection: b[p*nu+u] -> x[ix*ny+iy] (overwritten).
int host_backward(double* b, double* x, sgdct_projector* h) {
  if (!h || !b || !x) return fail(SGDCT_EINVAL, "null argument");
  try {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "set device");
    int rc = ensure_staging(h);
    if (rc) return rc;
    const int nx = h->nx, ny = h->ny;
    const size_t ni = (size_t)nx * ny, ns = (size_t)h->np * h->nu;
    for (size_t i = 0; i < ns; ++i) h->h_sino[i] = (float)b[i];
    e = hipMemcpy(h->d_sino, h->h_sino, ns * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "upload sinogram");
    rc = run_backward(h, h->d_sino, h->d_img, 1, nullptr);
    if (rc) return rc;
    e = hipMemcpy(h->h_img, h->d_img, ni * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "download image");
    for (int ix = 0; ix < nx; ++ix)
      for (int iy = 0; iy < ny; ++iy) x[(size_t)ix * ny + iy] = (double)h->h_img[(size_t)iy * nx + ix];
    return SGDCT_OK;
  } catch (...) {
    return fail(SGDCT_EDEVICE, "unexpected exception in backward");
  }
}

}  // namespace

extern "C" {

sgdct_projector* b0_init(int nx, int ny, int np, int nu, double dx, double dy, double du, double lsd,
                         double lso, const double* angles, const double* dz, const double* drho) {
  (void)dz;  // out-of-plane source offset: no effect on a 2-D fan (z = 0 slice)
  return create(BASIS_B0, nx, ny, np, nu, dx, dy, du, lsd, lso, angles, drho);
}
int b0_forward_projection(double* b, double* x, sgdct_projector* h) { return host_forward(b, x, h); }
int b0_backward_projection(double* b, double* x, sgdct_projector* h) { return host_backward(b, x, h); }

sgdct_projector* b1_init(int nx, int ny, int np, int nu, double dx, double dy, double du, double lsd,
                         double lso, const double* angles, const double* dz, const double* drho) {
  (void)dz;
  return create(BASIS_B1, nx, ny, np, nu, dx, dy, du, lsd, lso, angles, drho);
}
int b1_forward_projection(double* b, double* x, sgdct_projector* h) { return host_forward(b, x, h); }
int b1_backward_projection(double* b, double* x, sgdct_projector* h) { return host_backward(b, x, h); }

sgdct_projector* SF_init(int nx, int ny, int np, int nu, double dx, double dy, double du, double lsd,
                         double lso, const double* angles) {
  return create(BASIS_SF, nx, ny, np, nu, dx, dy, du, lsd, lso, angles, nullptr);
}
int SF_forward_projection(double* b, double* x, sgdct_projector* h) { return host_forward(b, x, h); }
int SF_backward_projection(double* b, double* x, sgdct_projector* h) { return host_backward(b, x, h); }

// layers/A_H0.hh:24 / A_H0.cu:509-643 and A_Hl.hh:24 / A_Hl.cu:521-657:
// nl rows x nc columns, lattice constant dt; host image x[ic*nl + il].
sgdct_projector* h0_init(int nl, int nc, int np, int nu, double dt, double du, double lsd, double lso,
                         const double* angles, const double* dz, const double* drho) {
  (void)dz;
  if (!(dt > 0)) {
    fail(SGDCT_EINVAL, "dt must be positive");
    return nullptr;
  }
  return create(BASIS_H0, nc, nl, np, nu, dt * std::sqrt(3.0), 1.5 * dt, du, lsd, lso, angles, drho, dt);
}
int h0_forward_projection(double* b, double* x, sgdct_projector* h) { return host_forward(b, x, h); }
int h0_backward_projection(double* b, double* x, sgdct_projector* h) { return host_backward(b, x, h); }

sgdct_projector* hl_init(int nl, int nc, int np, int nu, double dt, double du, double lsd, double lso,
                         const double* angles, const double* dz, const double* drho) {
  (void)dz;
  if (!(dt > 0)) {
    fail(SGDCT_EINVAL, "dt must be positive");
    return nullptr;
  }
  return create(BASIS_HL, nc, nl, np, nu, dt * std::sqrt(3.0), 0.5 * dt, du, lsd, lso, angles, drho, dt);
}
int hl_forward_projection(double* b, double* x, sgdct_projector* h) { return host_forward(b, x, h); }
int hl_backward_projection(double* b, double* x, sgdct_projector* h) { return host_backward(b, x, h); }

void sgdct_destroy(sgdct_projector* h) { delete h; }

int sgdct_forward_dev(sgdct_projector* h, const float* img, float* sino, int batch, void* stream) {
  if (!h || !img || !sino || batch <= 0) return fail(SGDCT_EINVAL, "bad forward_dev arguments");
  return run_forward(h, img, sino, batch, (hipStream_t)stream);
}

int sgdct_backward_dev(sgdct_projector* h, const float* sino, float* img, int batch, void* stream) {
  if (!h || !img || !sino || batch <= 0) return fail(SGDCT_EINVAL, "bad backward_dev arguments");
  return run_backward(h, sino, img, batch, (hipStream_t)stream);
}

int sgdct_shape(const sgdct_projector* h, int* out) {
  if (!h || !out) return fail(SGDCT_EINVAL, "null argument");
  out[0] = h->nx; out[1] = h->ny; out[2] = h->np; out[3] = h->nu;
  return h->basis;
}

long long sgdct_eval_count(sgdct_projector* h) {
  if (!h) return fail(SGDCT_EINVAL, "null handle");
  return eval_count(h);
}

int sgdct_tv_gradient(const float* x, float* g, int nx, int ny, int batch, float eps, void* stream) {
  if (!x || !g || nx <= 0 || ny <= 0 || batch <= 0 || !(eps > 0)) return fail(SGDCT_EINVAL, "bad tv_gradient arguments");
  hipError_t e = launch_tv_grad(x, g, nx, ny, batch, eps, (hipStream_t)stream);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "tv_gradient launch");
}

int sgdct_sirt_residual(const float* b, const float* ax, const float* w, float* r, long long n, long long per,
                        void* stream) {
  if (!b || !ax || !w || !r || n <= 0 || per <= 0 || n % per) return fail(SGDCT_EINVAL, "bad sirt_residual arguments");
  hipError_t e = launch_resid(b, ax, w, r, (size_t)n, (size_t)per, (hipStream_t)stream);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "sirt_residual launch");
}

int sgdct_sirt_update(float* x, const float* c, const float* v, float lam, int positivity, long long n, long long per,
                      void* stream) {
  if (!x || !c || !v || n <= 0 || per <= 0 || n % per) return fail(SGDCT_EINVAL, "bad sirt_update arguments");
  hipError_t e = launch_update(x, c, v, lam, positivity, (size_t)n, (size_t)per, (hipStream_t)stream);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "sirt_update launch");
}

int sgdct_axpy(float* x, const float* g, float a, long long n, void* stream) {
  if (!x || !g || n <= 0) return fail(SGDCT_EINVAL, "bad axpy arguments");
  hipError_t e = launch_axpy(x, g, a, (size_t)n, (hipStream_t)stream);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "axpy launch");
}

const char* sgdct_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
