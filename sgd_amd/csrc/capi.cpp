// capi.cpp -- C ABI (include/sgdct.h): handles, host-side FP64 geometry
// tables, host-array entry points with the reference's layouts, and FP32
// device-pointer entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sgdct.h"
#include "footprint.hpp"
#include "projector.hpp"

namespace sgdct {
hipError_t launch_forward(int basis, const Proj2DArgs& g, const float* img, float* sino, int batch, int nseg,
                          double* part, hipStream_t s);
hipError_t launch_backward(int basis, const Proj2DArgs& g, bool use_lds, const float* sino, float* img, int batch,
                           int ngroup, double* part, hipStream_t s);
int max_tile_window(const Proj2DArgs& g, const ViewGeo* views_host);
hipError_t launch_count(const Proj2DArgs& g, unsigned long long* out, hipStream_t s);
hipError_t launch_sf_forward(const Proj2DArgs& g, const float* img, float* sino, int batch, int nseg, double* part,
                             hipStream_t s);
hipError_t launch_sf_backward(const Proj2DArgs& g, const float* sino, float* img, int batch, int ngroup,
                              double* part, hipStream_t s);
void hex_bin_constants(int basis, const ViewGeo& v, int nu, double dt, void* out);
size_t hex_bin_size(int basis);
hipError_t launch_hex_forward(int basis, const Proj2DArgs& g, double dt, const float* img, float* sino, int batch,
                              int nseg, double* part, hipStream_t s);
hipError_t launch_hex_backward(int basis, const Proj2DArgs& g, double dt, const float* sino, float* img, int batch,
                               int ngroup, double* part, hipStream_t s);
hipError_t launch_hex_count(const Proj2DArgs& g, double dt, unsigned long long* out, hipStream_t s);
hipError_t launch_tv_grad(const float* x, float* g, int nx, int ny, int batch, float eps, hipStream_t s);
hipError_t launch_tv_descent(float* x, float* g, double* part, const double* step, int nx, int ny, int batch,
                             float eps, int nsteps, hipStream_t s);
hipError_t launch_resid(const float* b, const float* ax, const float* w, float* r, size_t n, size_t per, hipStream_t s);
hipError_t launch_update(float* x, const float* c, const float* v, float lam, int positivity, size_t n, size_t per,
                         hipStream_t s);
hipError_t launch_axpy(float* x, const float* g, float a, size_t n, hipStream_t s);
hipError_t launch_transpose(const float* in, float* out, int rows, int cols, int batch, hipStream_t s);
hipError_t launch_cone_forward_chunk(int model, const Cone3DArgs& a, int p0, int nview, const float* xT, float* Y,
                                     float* proj, hipStream_t s);
hipError_t launch_cone_backward_chunk(int model, const Cone3DArgs& a, int p0, int nview, const float* projT,
                                      float* Yb, float* xT, int accumulate, hipStream_t s);
hipError_t launch_matgen(int basis, const Proj2DArgs& g, bool write, unsigned int* counts, const long long* offsets,
                         double* out, hipStream_t s);
hipError_t launch_hex_matgen(int basis, const Proj2DArgs& g, double dt, bool write, unsigned int* counts,
                             const long long* offsets, double* out, hipStream_t s);
hipError_t launch_batch_forward(const BatchPlan& P, int blocks, const float* img, float* sino, hipStream_t s);
hipError_t nn_conv3x3(const void* in, const void* in2, int c0, const void* w, const float* bias, void* out,
                      const void* mask, int B, int H, int W, int cin, int cout, int relu, int out_f32, hipStream_t s);
long long nn_wgrad_splits(long long M);
hipError_t nn_wgrad(const void* x, const void* x2, int c0, const void* dy, float* dw, float* db, float* partial, int B,
                    int H, int W, int cin, int cout, hipStream_t s);
hipError_t nn_colsum_bf16(const void* a, float* out, float* partial, long long M, int C, hipStream_t s);
hipError_t nn_pad_input(const float* x, void* out8, long long n, hipStream_t s);
hipError_t nn_pool2(const void* src, void* dst, int B, int H, int W, int C, hipStream_t s);
hipError_t nn_pool2_back(const float* dv, float* gsrc, void* out, const void* mask, int B, int H, int W, int C,
                         hipStream_t s);
hipError_t nn_upcat(const void* low, const void* skip, void* dst, int B, int H, int W, int C0, int C1,
                    hipStream_t s);
hipError_t nn_upcat_back(const float* dv, float* glow, void* glow_out, const void* glow_mask, float* gskip,
                         int skip_accumulate, int B, int H, int W, int C0, int C1, hipStream_t s);
hipError_t nn_head(const void* d0, const float* w, const float* b, const float* x, float* y, long long M, int C,
                   hipStream_t s);
hipError_t nn_head_back(const float* y, const float* target, const float* w, const void* d0, float* dy, float* gd0,
                        void* gd0_out, const void* mask, float* gw, float* partial, long long M, int C,
                        hipStream_t s);
hipError_t nn_relu_mask(const float* g, const void* act, void* out, long long n, hipStream_t s);
hipError_t nn_flip_weights(const void* w, void* wt, int cin, int cout, hipStream_t s);
hipError_t nn_adam(float* w32, void* w16, const float* g, float* m, float* v, long long n, float lr, float b1,
                   float b2, float eps, int step, hipStream_t s);
hipError_t nn_fill(float* p, long long n, float value, hipStream_t s);
hipError_t nn_to_bf16(const float* src, void* dst, long long n, hipStream_t s);
hipError_t launch_axpy_batch(float* x, const float* g, const float* a, size_t n, size_t per, hipStream_t s);
hipError_t launch_batch_backward(const BatchPlan& P, int blocks, const float* sino, float* img, hipStream_t s);
}  // namespace sgdct

using namespace sgdct;

// Partial sums of sgdct_project_dev_noalloc (FP64, one array per GPU, alive
// until the process ends).  Host code never passes `scratch` itself to a
// kernel -- on the host it names a stand-in, not device memory -- but the
// device address scratch_on() reads with hipGetSymbolAddress.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// The failing call also left `e` as the runtime's last error, which the next
// launch check (hipGetLastError after a kernel launch) would report as its
// own; it is taken here so that a failed allocation does not fail the next
// call as well.
int hip_fail(hipError_t e, const char* what) {
  (void)hipGetLastError();
  return fail(e == hipErrorOutOfMemory ? SGDCT_ENOMEM : SGDCT_EDEVICE,
              std::string(what) + ": " + hipGetErrorString(e));
}

// Device address of `scratch` on `device` (which must be current), looked up
// once per device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_on(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(SGDCT_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "scratch array address");
    if (!p) return fail(SGDCT_EDEVICE, "scratch array address is null");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return SGDCT_OK;
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
  // B0 / B1 / SF: A splits each bin's walk into line segments and A^T the views
  // into groups; fwd_auto / bwd_auto are chosen at creation from one image's
  // launch size, fwd_split / bwd_split (0 = automatic) override them.  The
  // FP64 partial sums go to d_part (part_cap doubles, allocated at creation
  // or by sgdct_split_config, never inside a projection call); a batch is
  // processed in as many launches as the workspace requires, so every image
  // is computed the same way whatever the batch size.
  int fwd_split = 0, bwd_split = 0, fwd_auto = 1, bwd_auto = 1;
  double* d_part = nullptr;
  size_t part_cap = 0;
  Proj2DArgs args{};
  // cone beam (BASIS_B1_3D): volume nz slices, detector nv rows
  int nz = 1, nv = 1;
  double dz = 0, dv = 0;
  int cone_model = 0;         // 0 B1, 1 SF-TR, 2 SF-TT
  int chunk = 0;              // views per pass of the staged kernels
  float* d_xT = nullptr;      // [ncol][nz] column-major volume
  float* d_Y = nullptr;       // [chunk][ncol][nv] per-column row profiles
  float* d_projT = nullptr;   // [chunk][nu][nv]
  // A's passes alternate between two row-profile buffers and two streams
  // (the caller's and cone_s2), so that one pass's in-plane kernel runs
  // beside the next pass's axial kernel: B1 A 55.0 -> 51.8 ms at 256^3, 360
  // views, 512 x 384 (16 views per pass).  The same overlap of A^T's passes
  // (in-plane kernels on the caller's stream, volume updates in pass order on
  // cone_s2) gained nothing (39.0 -> 39.9 ms) and is not used.
  float* d_Y2 = nullptr;
  hipStream_t cone_s2 = nullptr;
  hipEvent_t cone_ev[2] = {nullptr, nullptr};

  // Frees the cone-beam workspace and clears its pointers: the workspace is
  // either complete or absent, also after an allocation that failed halfway.
  void release_cone() {
    for (float** b : {&d_xT, &d_Y, &d_projT, &d_Y2}) {
      if (*b) (void)hipFree(*b);
      *b = nullptr;
    }
    if (cone_s2) (void)hipStreamDestroy(cone_s2);
    cone_s2 = nullptr;
    for (hipEvent_t& ev : cone_ev) {
      if (ev) (void)hipEventDestroy(ev);
      ev = nullptr;
    }
  }

  ~sgdct_projector() {
    if (d_views) (void)hipFree(d_views);
    if (d_bins) (void)hipFree(d_bins);
    if (d_img) (void)hipFree(d_img);
    if (d_sino) (void)hipFree(d_sino);
    if (d_count) (void)hipFree(d_count);
    if (d_part) (void)hipFree(d_part);
    release_cone();
    if (h_img) (void)hipHostFree(h_img);
    if (h_sino) (void)hipHostFree(h_sino);
  }
};

struct sgdct_batch {
  int n = 0, device = 0, ncu = 256;
  std::vector<long long> img_off, sino_off;  // [n + 1]
  std::vector<SliceDesc> desc;
  std::vector<sgdct_projector*> proj;        // the slices' handles (per-slice mode)
  std::vector<double> fwd_cost, bwd_cost;    // per-item cost estimates
  std::vector<long long> fwd_items, bwd_items;
  SliceDesc* d_slices = nullptr;
  int* d_order = nullptr;          // [2][n]
  long long* d_prefix = nullptr;   // [2][n + 1]
  unsigned int* d_counters = nullptr;
  BatchPlan plan{};
  // Persistent blocks per CU of A and A^T; 0 = one launch per slice (each
  // projector's own split kernels).  Defaults: the autotuner's persistent
  // choice on 256 uneven slices (A^T at 8 blocks per CU: 165.5 ms against
  // 200.8 ms at 4).
  int fwd_bpc = 8, bwd_bpc = 8, order = 0;
  struct Graph {
    hipGraphExec_t exec = nullptr;
    const void* in = nullptr;
    void* out = nullptr;
  } graph[2];

  void drop_graphs() {
    for (auto& g : graph) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = Graph{};
    }
  }
  ~sgdct_batch() {
    drop_graphs();
    if (d_slices) (void)hipFree(d_slices);
    if (d_order) (void)hipFree(d_order);
    if (d_prefix) (void)hipFree(d_prefix);
    if (d_counters) (void)hipFree(d_counters);
  }
};

namespace {

int auto_fwd_split(const sgdct_projector* h);
int auto_bwd_split(const sgdct_projector* h);
int ensure_part(sgdct_projector* h, size_t doubles);
void release_part(sgdct_projector* h);
size_t part_need(const sgdct_projector* h);
int part_images(const sgdct_projector* h);

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
      // constants in the offset t' = s t scaled to the support (footprint.hpp)
      B1Bin r{};
      const double a = std::max(vx, vy);
      const double b = std::min(vx, vy);
      double sc, scale;
      if (b > 1e-7) {
        sc = 1.0 / (a + b);
        const double as = a * sc, bs = b * sc;
        const double h = 12.0 * as * as * bs * bs;  // 1 / (2 c')
        r.h = (float)h;
        r.b = (float)bs;
        r.a = (float)as;
        r.nk2 = (float)(-12.0 * bs * bs);
        r.hk4 = (float)(h - 2.0 * bs * bs * bs * bs);
        scale = sc / (2.0 * h);  // s c'
      } else {  // 1tri1box limit, spline_conv.hh:393-405
        sc = 1.0 / a;
        r.h = 0.5f;
        r.b = 0.0f;
        r.a = 1.0f;
        r.nk2 = -0.5f;
        r.hk4 = 0.5f;
        scale = sc;
      }
      r.ka = (float)(ka * sc);
      r.kb = (float)(kb * sc);
      r.scale = (float)scale;
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
// fan-beam handles split both A (line or row segments) and A^T (view groups)
bool splits(int basis) { return basis == BASIS_B0 || basis == BASIS_B1 || basis == BASIS_SF || is_hex(basis); }

// ushift moves the bin edges by ushift bins: 0.5 gives cells centred on
// integer detector coordinates (layers/SF_projector.cu:209-216).
sgdct_projector* create(int basis, int nx, int ny, int np, int nu, double dx, double dy, double du,
                        double lsd, double lso, const double* angles, const double* drho, double dt = 0.0,
                        double ushift = 0.0) {
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
    v.e0x = dtv[0] - (0.5 * nu + ushift) * pu[0] - src[0];
    v.e0y = dtv[1] - (0.5 * nu + ushift) * pu[1] - src[1];
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
  if (splits(basis)) {
    // hexagonal A runs unsplit unless asked: at 296 x 256 points, 1024 views
    // and 1024 bins, 2..16 row segments took one image from 1.11 to 1.20..1.40
    // ms (H0) and a batch of 64 from 59 to 64..84 ms (measured when A walked
    // lattice rows only, before steep waves walked the vertical lines)
    h->fwd_auto = is_hex(basis) ? 1 : auto_fwd_split(h);
    h->bwd_auto = auto_bwd_split(h);
    if (ensure_part(h, part_images(h) * part_need(h)) != SGDCT_OK) {
      delete h;
      return nullptr;
    }
  }
  return h;
}

size_t img_elems(const sgdct_projector* h) { return (size_t)h->nx * h->ny * h->nz; }
size_t sino_elems(const sgdct_projector* h) { return (size_t)h->np * h->nu * h->nv; }

int ensure_staging(sgdct_projector* h) {
  if (h->d_img) return SGDCT_OK;
  const size_t ni = img_elems(h), ns = sino_elems(h);
  hipError_t e = hipMalloc(&h->d_img, ni * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&h->d_sino, ns * sizeof(float));
  if (e == hipSuccess) e = hipHostMalloc(&h->h_img, ni * sizeof(float));
  if (e == hipSuccess) e = hipHostMalloc(&h->h_sino, ns * sizeof(float));
  if (e != hipSuccess) {
    // all four or none: the guard above tests d_img only
    for (float** b : {&h->d_img, &h->d_sino}) {
      if (*b) (void)hipFree(*b);
      *b = nullptr;
    }
    for (float** b : {&h->h_img, &h->h_sino}) {
      if (*b) (void)hipHostFree(*b);
      *b = nullptr;
    }
    return hip_fail(e, "staging allocation");
  }
  return SGDCT_OK;
}

Cone3DArgs cone_args(const sgdct_projector* h) {
  Cone3DArgs a;
  a.g = h->args;
  a.nz = h->nz;
  a.nv = h->nv;
  a.nrc = (h->nv + 63) / 64;
  a.vc = (float)(0.5 * h->nv);
  a.ldz = h->lsd * h->dz / h->dv;
  a.cvol = (float)(h->dx * h->dy * h->dz);
  a.dz = (float)h->dz;
  a.du = h->du;
  a.fac = (float)(h->lsd * h->lsd * h->dx * h->dy * h->dz / (h->du * h->dv));
  return a;
}

bool is_cone(int basis) { return basis == BASIS_B1_3D || basis == BASIS_SF_3D; }

int ensure_cone_workspace(sgdct_projector* h) {
  if (h->d_xT) return SGDCT_OK;
  const size_t ncol = (size_t)h->nx * h->ny;
  const size_t per_view = ncol * h->nv * sizeof(float);
  // views per pass: as many as a 2 GiB row-profile buffer holds (16 at 256^2
  // columns x 384 rows), a multiple of 8 so that every XCD takes whole views
  size_t chunk = std::max<size_t>(1, ((size_t)2 << 30) / per_view);
  if (chunk >= 8) chunk -= chunk % 8;
  h->chunk = (int)std::min<size_t>(chunk, (size_t)h->np);
  hipError_t e = hipMalloc(&h->d_xT, ncol * h->nz * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&h->d_Y, (size_t)h->chunk * per_view);
  if (e == hipSuccess) e = hipMalloc(&h->d_projT, (size_t)h->chunk * h->nu * h->nv * sizeof(float));
  if (e == hipSuccess && h->np > h->chunk) {
    e = hipMalloc(&h->d_Y2, (size_t)h->chunk * per_view);
    // non-blocking: no implicit ordering against the null stream; every
    // dependency is an explicit event
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->cone_s2, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&h->cone_ev[k], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    h->release_cone();
    return hip_fail(e, "cone-beam workspace");
  }
  return SGDCT_OK;
}

// Volume [nz][ny][nx] -> projections [np][nv][u], `batch` back to back.
int cone_forward(sgdct_projector* h, const float* vol, float* proj, int batch, hipStream_t s) {
  int rc = ensure_cone_workspace(h);
  if (rc) return rc;
  const Cone3DArgs a = cone_args(h);
  const int ncol = h->nx * h->ny;
  for (int b = 0; b < batch; ++b) {
    const float* x = vol + (size_t)b * ncol * h->nz;
    float* y = proj + (size_t)b * h->np * h->nv * h->nu;
    hipError_t e = launch_transpose(x, h->d_xT, h->nz, ncol, 1, s);
    // odd passes on cone_s2 with d_Y2, after the transpose; the caller's
    // stream waits for them before the next image's transpose and at the end
    const bool two = h->cone_s2 != nullptr;
    if (two && e == hipSuccess) e = hipEventRecord(h->cone_ev[0], s);
    if (two && e == hipSuccess) e = hipStreamWaitEvent(h->cone_s2, h->cone_ev[0], 0);
    for (int p0 = 0, k = 0; p0 < h->np && e == hipSuccess; p0 += h->chunk, ++k) {
      const bool odd = two && (k & 1);
      e = launch_cone_forward_chunk(h->cone_model, a, p0, std::min(h->chunk, h->np - p0), h->d_xT,
                                    odd ? h->d_Y2 : h->d_Y, y, odd ? h->cone_s2 : s);
    }
    if (two) {
      // joined also after a failed launch: no pass of this call stays
      // unordered against the caller's stream
      hipError_t j = hipEventRecord(h->cone_ev[1], h->cone_s2);
      if (j == hipSuccess) j = hipStreamWaitEvent(s, h->cone_ev[1], 0);
      if (e == hipSuccess) e = j;
    }
    if (e != hipSuccess) return hip_fail(e, "cone-beam forward launch");
  }
  return SGDCT_OK;
}

int cone_backward(sgdct_projector* h, const float* proj, float* vol, int batch, hipStream_t s) {
  int rc = ensure_cone_workspace(h);
  if (rc) return rc;
  const Cone3DArgs a = cone_args(h);
  const int ncol = h->nx * h->ny;
  for (int b = 0; b < batch; ++b) {
    const float* y = proj + (size_t)b * h->np * h->nv * h->nu;
    float* x = vol + (size_t)b * ncol * h->nz;
    hipError_t e = hipSuccess;
    for (int p0 = 0; p0 < h->np && e == hipSuccess; p0 += h->chunk) {
      const int nview = std::min(h->chunk, h->np - p0);
      e = launch_transpose(y + (size_t)p0 * h->nv * h->nu, h->d_projT, h->nv, h->nu, nview, s);
      if (e == hipSuccess)
        e = launch_cone_backward_chunk(h->cone_model, a, p0, nview, h->d_projT, h->d_Y, h->d_xT, p0 > 0, s);
    }
    if (e == hipSuccess) e = launch_transpose(h->d_xT, x, ncol, h->nz, 1, s);
    if (e != hipSuccess) return hip_fail(e, "cone-beam backward launch");
  }
  return SGDCT_OK;
}

// Launch-size targets of the automatic splits (one image): enough waves for
// the GPU's 8192 wave slots (256 CUs x 4 SIMDs x 8) to turn over about 24
// times in A (its per-wave work is uneven) and 12 times in A^T.  B1 at
// 512 x 512, 720 views, 768 bins thus runs 24 segments and 24 view groups:
// A 0.942 ms, A^T 0.922 ms, against 0.960 / 0.932 ms at 16 / 16.
constexpr long long kFwdWaveTarget = 200000, kBwdWaveTarget = 98304;
constexpr int kAutoMaxSplit = 24;  // automatic splits stop here
constexpr int kMaxSplit = 32;      // sgdct_split_config accepts 1..kMaxSplit

int auto_fwd_split(const sgdct_projector* h) {
  const long long waves = (long long)h->np * h->args.nchunk;
  return (int)std::max(1LL, std::min<long long>(kAutoMaxSplit, (kFwdWaveTarget + waves - 1) / waves));
}

int auto_bwd_split(const sgdct_projector* h) {
  const long long waves = 4LL * ((h->nx + kTile - 1) / kTile) * ((h->ny + kTile - 1) / kTile);
  const long long by_views = std::max(1, h->np / (4 * kViewChunk));  // groups of >= 16 views
  return (int)std::max(1LL, std::min<long long>(std::min<long long>(kAutoMaxSplit, by_views),
                                                (kBwdWaveTarget + waves - 1) / waves));
}

// Room for `doubles` partial sums.
int ensure_part(sgdct_projector* h, size_t doubles) {
  if (doubles <= h->part_cap) return SGDCT_OK;
  if (h->d_part) (void)hipFree(h->d_part);
  h->d_part = nullptr;
  h->part_cap = 0;
  hipError_t e = hipMalloc(&h->d_part, doubles * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "partial-sum workspace");
  h->part_cap = doubles;
  return SGDCT_OK;
}

void release_part(sgdct_projector* h) {
  if (h->d_part) (void)hipFree(h->d_part);
  h->d_part = nullptr;
  h->part_cap = 0;
  h->fwd_auto = h->bwd_auto = 1;
}

// Workspace of one image at the current splits.
size_t part_need(const sgdct_projector* h) {
  const int fs = h->fwd_split ? h->fwd_split : h->fwd_auto;
  const int bs = h->bwd_split ? h->bwd_split : h->bwd_auto;
  return std::max(fs > 1 ? (size_t)fs * h->np * h->nu : 0, bs > 1 ? (size_t)bs * h->nx * h->ny : 0);
}

// Images per split launch: at least 2, and more (up to 16) while the
// workspace stays within kPartBytes.  A batch runs as launches of this many
// images; launches of few small images lose a larger share to their tails
// (hexagonal A^T at 296 x 256 points, batch 64: 2 images per launch 60.9 ms,
// 8 images 58.7 ms).
constexpr size_t kPartBytes = (size_t)64 << 20;

int part_images(const sgdct_projector* h) {
  const size_t per = part_need(h) * sizeof(double);
  if (per == 0) return 2;
  return (int)std::max<size_t>(2, std::min<size_t>(16, kPartBytes / per));
}


hipError_t launch_fwd2d(const sgdct_projector* h, const float* img, float* sino, int batch, int nseg, double* part,
                        hipStream_t s) {
  if (h->basis == BASIS_SF) return launch_sf_forward(h->args, img, sino, batch, nseg, part, s);
  if (is_hex(h->basis)) return launch_hex_forward(h->basis, h->args, h->dt, img, sino, batch, nseg, part, s);
  return launch_forward(h->basis, h->args, img, sino, batch, nseg, part, s);
}

hipError_t launch_bwd2d(const sgdct_projector* h, const float* sino, float* img, int batch, int ngroup, double* part,
                        hipStream_t s) {
  if (h->basis == BASIS_SF) return launch_sf_backward(h->args, sino, img, batch, ngroup, part, s);
  if (is_hex(h->basis)) return launch_hex_backward(h->basis, h->args, h->dt, sino, img, batch, ngroup, part, s);
  return launch_backward(h->basis, h->args, h->bwd_lds, sino, img, batch, ngroup, part, s);
}

// Fan-beam A of `batch` images with each bin's walk in `nseg` line segments,
// the FP64 partial sums in part[0, cap) (doubles): as many images per launch
// as the buffer holds, so every image is computed the same way whatever the
// batch size or the buffer.
int fan_forward(sgdct_projector* h, const float* img, float* sino, int batch, int nseg, double* part, size_t cap,
                hipStream_t s) {
  hipError_t e = hipSuccess;
  const size_t ni = (size_t)h->nx * h->ny, ns = (size_t)h->np * h->nu, per = (size_t)nseg * ns;
  if (nseg == 1) {
    e = launch_fwd2d(h, img, sino, batch, 1, nullptr, s);
  } else {
    if (!part || cap < per) return fail(SGDCT_EDEVICE, "partial-sum workspace missing");
    const int bc = (int)std::min<size_t>(batch, cap / per);
    for (int b0 = 0; b0 < batch && e == hipSuccess; b0 += bc)
      e = launch_fwd2d(h, img + b0 * ni, sino + b0 * ns, std::min(bc, batch - b0), nseg, part, s);
  }
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "forward launch");
}

// Fan-beam A^T with the views in `ngroup` groups, partial sums as above.
int fan_backward(sgdct_projector* h, const float* sino, float* img, int batch, int ngroup, double* part, size_t cap,
                 hipStream_t s) {
  hipError_t e = hipSuccess;
  const size_t ni = (size_t)h->nx * h->ny, ns = (size_t)h->np * h->nu, per = (size_t)ngroup * ni;
  if (ngroup == 1) {
    e = launch_bwd2d(h, sino, img, batch, 1, nullptr, s);
  } else {
    if (!part || cap < per) return fail(SGDCT_EDEVICE, "partial-sum workspace missing");
    const int bc = (int)std::min<size_t>(batch, cap / per);
    for (int b0 = 0; b0 < batch && e == hipSuccess; b0 += bc)
      e = launch_bwd2d(h, sino + b0 * ns, img + b0 * ni, std::min(bc, batch - b0), ngroup, part, s);
  }
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "backward launch");
}

int fwd_segments(const sgdct_projector* h) { return h->fwd_split ? h->fwd_split : h->fwd_auto; }
int bwd_groups(const sgdct_projector* h) { return h->bwd_split ? h->bwd_split : h->bwd_auto; }

int run_forward(sgdct_projector* h, const float* img, float* sino, int batch, hipStream_t s) {
  if (is_cone(h->basis)) return cone_forward(h, img, sino, batch, s);
  return fan_forward(h, img, sino, batch, fwd_segments(h), h->d_part, h->part_cap, s);
}

int run_backward(sgdct_projector* h, const float* sino, float* img, int batch, hipStream_t s) {
  if (is_cone(h->basis)) return cone_backward(h, sino, img, batch, s);
  return fan_backward(h, sino, img, batch, bwd_groups(h), h->d_part, h->part_cap, s);
}

// Footprint evaluations of the reference's cone-beam forward
// (A_B1_3D.cu:261 runs once per bin of the column's clamped corner range for
// every slice before the first whose axial trapezoid starts past the last
// row, A_B1_3D.cu:204).  Host FP64, from the handle's geometry.
long long cone_eval_count(const sgdct_projector* h, const std::vector<ViewGeo>& views) {
  long long total = 0;
  const double vc = 0.5 * h->nv, ldz = h->lsd * h->dz / h->dv, zoff = 0.5 * (h->nz - 1);
  for (int p = 0; p < h->np; ++p) {
    const ViewGeo& v = views[p];
    for (int iy = 0; iy < h->ny; ++iy)
      for (int ix = 0; ix < h->nx; ++ix) {
        double umin = 1e300, umax = -1e300;
        for (int k = 0; k < 4; ++k) {
          const double wx = (ix + ((k & 1) ? 1.0 : -1.0)) * h->dx + v.w0x;
          const double wy = (iy + ((k & 2) ? 1.0 : -1.0)) * h->dy + v.w0y;
          const double uu = (v.e0x * wy - v.e0y * wx) / std::fabs(v.pux * wy - v.puy * wx);
          umin = std::min(umin, uu);
          umax = std::max(umax, uu);
        }
        const int lo = std::min(std::max(0, (int)std::floor(umin)), h->nu - 1);
        const int hi = std::min(std::max(0, (int)std::floor(umax)), h->nu - 1);
        const double wy = iy * h->dy + v.w0y;
        const double da = -(v.pux * wy - v.puy * ((ix - 1) * h->dx + v.w0x)) / h->du;
        const double db = -(v.pux * wy - v.puy * ((ix + 1) * h->dx + v.w0x)) / h->du;
        const double kmin = std::min(ldz / da, ldz / db);
        // first slice with vc + kmin (iz - zoff - 1/2) >= nv
        const double first = std::ceil((h->nv - vc) / kmin + zoff + 0.5);
        const long long nzeff = (long long)std::min<double>(std::max(first, 0.0), (double)h->nz);
        total += (long long)(hi - lo + 1) * nzeff;
      }
  }
  return total;
}

long long eval_count(sgdct_projector* h) {
  if (h->evals >= 0) return h->evals;
  if (h->basis == BASIS_SF_3D) return fail(SGDCT_EUNSUP, "no evaluation count for the separable-footprint projector");
  if (h->basis == BASIS_B1_3D) {
    std::vector<ViewGeo> views(h->np);
    hipError_t e = hipMemcpy(views.data(), h->d_views, sizeof(ViewGeo) * h->np, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "evaluation count");
    h->evals = cone_eval_count(h, views);
    return h->evals;
  }
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
// The host image layout x[ix * ny + iy] to the device's [iy][ix] (FP64 ->
// FP32) and back, in 32 x 32 blocks: a plain double loop writes (or reads)
// one element per cache line along the strided side, and on the 512 x 512
// workload that took ~1 ms of each host-array call.
constexpr int kHostBlock = 32;

void host_to_rows(const double* src, float* dst, int nx, int ny) {
  for (int x0 = 0; x0 < nx; x0 += kHostBlock)
    for (int y0 = 0; y0 < ny; y0 += kHostBlock) {
      const int x1 = std::min(nx, x0 + kHostBlock), y1 = std::min(ny, y0 + kHostBlock);
      for (int iy = y0; iy < y1; ++iy)
        for (int ix = x0; ix < x1; ++ix) dst[(size_t)iy * nx + ix] = (float)src[(size_t)ix * ny + iy];
    }
}

void rows_to_host(const float* src, double* dst, int nx, int ny) {
  for (int x0 = 0; x0 < nx; x0 += kHostBlock)
    for (int y0 = 0; y0 < ny; y0 += kHostBlock) {
      const int x1 = std::min(nx, x0 + kHostBlock), y1 = std::min(ny, y0 + kHostBlock);
      for (int ix = x0; ix < x1; ++ix)
        for (int iy = y0; iy < y1; ++iy) dst[(size_t)ix * ny + iy] = (double)src[(size_t)iy * nx + ix];
    }
}

int host_forward(double* b, double* x, sgdct_projector* h) {
  if (!h || !b || !x) return fail(SGDCT_EINVAL, "null argument");
  if (is_cone(h->basis)) return fail(SGDCT_EINVAL, "cone-beam handle: use the *_3d entry points");
  try {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "set device");
    int rc = ensure_staging(h);
    if (rc) return rc;
    const int nx = h->nx, ny = h->ny;
    host_to_rows(x, h->h_img, nx, ny);
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
  if (is_cone(h->basis)) return fail(SGDCT_EINVAL, "cone-beam handle: use the *_3d entry points");
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
    rows_to_host(h->h_img, x, nx, ny);
    return SGDCT_OK;
  } catch (...) {
    return fail(SGDCT_EDEVICE, "unexpected exception in backward");
  }
}

// Cone-beam host arrays: x copied raw ([iz][iy][ix], A_B1_3D.cu:716),
// b = [p][v][u] (A_B1_3D.cu:731).
long long cone_host_forward(double* b, double* x, sgdct_projector* h, int basis) {
  if (!h || !b || !x) return fail(SGDCT_EINVAL, "null argument");
  if (h->basis != basis) return fail(SGDCT_EINVAL, "handle of another projector");
  try {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "set device");
    int rc = ensure_staging(h);
    if (rc) return rc;
    const size_t ni = img_elems(h), ns = sino_elems(h);
    for (size_t i = 0; i < ni; ++i) h->h_img[i] = (float)x[i];
    e = hipMemcpy(h->d_img, h->h_img, ni * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "upload volume");
    rc = run_forward(h, h->d_img, h->d_sino, 1, nullptr);
    if (rc) return rc;
    e = hipMemcpy(h->h_sino, h->d_sino, ns * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "download projections");
    for (size_t i = 0; i < ns; ++i) b[i] = (double)h->h_sino[i];
    return basis == BASIS_B1_3D ? eval_count(h) : SGDCT_OK;
  } catch (...) {
    return fail(SGDCT_EDEVICE, "unexpected exception in forward");
  }
}

int cone_host_backward(double* b, double* x, sgdct_projector* h, int basis) {
  if (!h || !b || !x) return fail(SGDCT_EINVAL, "null argument");
  if (h->basis != basis) return fail(SGDCT_EINVAL, "handle of another projector");
  try {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "set device");
    int rc = ensure_staging(h);
    if (rc) return rc;
    const size_t ni = img_elems(h), ns = sino_elems(h);
    for (size_t i = 0; i < ns; ++i) h->h_sino[i] = (float)b[i];
    e = hipMemcpy(h->d_sino, h->h_sino, ns * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "upload projections");
    rc = run_backward(h, h->d_sino, h->d_img, 1, nullptr);
    if (rc) return rc;
    e = hipMemcpy(h->h_img, h->d_img, ni * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "download volume");
    for (size_t i = 0; i < ni; ++i) x[i] = (double)h->h_img[i];
    return SGDCT_OK;
  } catch (...) {
    return fail(SGDCT_EDEVICE, "unexpected exception in backward");
  }
}

// Relative per-item costs of the batched work queue: an A item is one view's
// 64-bin strip walk (length ~ the larger image edge), an A^T item one pixel
// tile over all views; per-evaluation weights B1 1, SF 0.6, B0 0.45.
double basis_weight(int basis) { return basis == BASIS_B1 ? 1.0 : (basis == BASIS_SF ? 0.6 : 0.45); }

// Upload the slice order and cumulative item counts for the current `order`
// (0: decreasing per-item cost, 1: as given).
int batch_upload_plan(sgdct_batch* b) {
  const int n = b->n;
  std::vector<int> ord[2];
  std::vector<long long> pre[2];
  for (int dir = 0; dir < 2; ++dir) {
    ord[dir].resize(n);
    for (int i = 0; i < n; ++i) ord[dir][i] = i;
    const std::vector<double>& cost = dir == 0 ? b->fwd_cost : b->bwd_cost;
    if (b->order == 0)
      std::stable_sort(ord[dir].begin(), ord[dir].end(), [&](int a, int c) { return cost[a] > cost[c]; });
    pre[dir].assign(n + 1, 0);
    const std::vector<long long>& items = dir == 0 ? b->fwd_items : b->bwd_items;
    for (int k = 0; k < n; ++k) pre[dir][k + 1] = pre[dir][k] + items[ord[dir][k]];
  }
  hipError_t e = hipMemcpy(b->d_order, ord[0].data(), sizeof(int) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_order + n, ord[1].data(), sizeof(int) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_prefix, pre[0].data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_prefix + n + 1, pre[1].data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "batch plan upload");
  BatchPlan& P = b->plan;
  P.slices = b->d_slices;
  P.fwd_order = b->d_order;
  P.bwd_order = b->d_order + n;
  P.fwd_prefix = b->d_prefix;
  P.bwd_prefix = b->d_prefix + n + 1;
  P.counters = b->d_counters;
  P.nslice = n;
  P.nfwd = pre[0][n];
  P.nbwd = pre[1][n];
  return SGDCT_OK;
}

int batch_run(sgdct_batch* b, int dir, const float* in, float* out, hipStream_t s) {
  if ((dir == 0 ? b->fwd_bpc : b->bwd_bpc) == 0) {
    // one launch per slice, in slice order, on the slices' own handles
    for (int i = 0; i < b->n; ++i) {
      sgdct_projector* h = b->proj[i];
      const int rc = dir == 0 ? run_forward(h, in + b->img_off[i], out + b->sino_off[i], 1, s)
                              : run_backward(h, in + b->sino_off[i], out + b->img_off[i], 1, s);
      if (rc) return rc;
    }
    return SGDCT_OK;
  }
  hipError_t e = dir == 0 ? launch_batch_forward(b->plan, b->ncu * b->fwd_bpc, in, out, s)
                          : launch_batch_backward(b->plan, b->ncu * b->bwd_bpc, in, out, s);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, dir == 0 ? "batch forward launch" : "batch backward launch");
}

// Explicit system matrix of B0 / B1 (pixel_matrix_generation,
// bilinear_matrix_generation): views from geo_init_example(lsd, lso, 0,
// 2 pi (np-1)/np) (utils/geo.cu:209-230), COO triplets in the reference's
// order.  Returns the triplet count or a negative code; a buffer smaller than
// the count is an error (the reference asserts, A_B1.cu:578-579).
int matrix_generation(int basis, int np, int nu, int nx, int ny, double dx, double dy, double du, double lsd,
                      double lso, double* spmat, int buffer_size, double dt = 0.0) {
  const bool query = !spmat && buffer_size == 0;  // count only
  if ((!spmat && !query) || buffer_size < 0) return fail(SGDCT_EINVAL, "null or negative matrix buffer");
  if (np <= 0) return fail(SGDCT_EINVAL, "np must be positive");
  std::vector<double> angles(np);
  const double end = M_PI * 2 * (np - 1) / np;
  const double dangle = np > 1 ? end / (np - 1) : 0.0;
  for (int p = 0; p < np; ++p) angles[p] = p * dangle;
  sgdct_projector* h = create(basis, nx, ny, np, nu, dx, dy, du, lsd, lso, angles.data(), nullptr, dt);
  if (!h) return SGDCT_EINVAL;
  auto launch = [&](bool write, unsigned int* counts, const long long* off, double* out) {
    return is_hex(basis) ? launch_hex_matgen(basis, h->args, dt, write, counts, off, out, nullptr)
                         : launch_matgen(basis, h->args, write, counts, off, out, nullptr);
  };
  const size_t nb = (size_t)np * nu;
  unsigned int* d_counts = nullptr;
  long long* d_off = nullptr;
  double* d_out = nullptr;
  int rc = SGDCT_OK;
  hipError_t e = hipMalloc(&d_counts, nb * sizeof(unsigned int));
  if (e == hipSuccess) e = launch(false, d_counts, nullptr, nullptr);
  std::vector<unsigned int> counts(nb);
  if (e == hipSuccess) e = hipMemcpy(counts.data(), d_counts, nb * sizeof(unsigned int), hipMemcpyDeviceToHost);
  long long total = 0;
  std::vector<long long> off(nb);
  if (e == hipSuccess) {
    for (size_t i = 0; i < nb; ++i) {
      off[i] = total;
      total += counts[i];
    }
    if (total > (long long)INT_MAX) {
      rc = fail(SGDCT_EUNSUP, "matrix has " + std::to_string(total) + " triplets, more than an int counts");
    } else if (query) {
      rc = SGDCT_OK;
    } else if (total > (long long)buffer_size) {
      rc = fail(SGDCT_EINVAL, "matrix buffer holds " + std::to_string(buffer_size) + " triplets, " +
                                  std::to_string(total) + " needed");
    } else if (total > 0) {
      e = hipMalloc(&d_off, nb * sizeof(long long));
      if (e == hipSuccess) e = hipMemcpy(d_off, off.data(), nb * sizeof(long long), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMalloc(&d_out, (size_t)total * 3 * sizeof(double));
      if (e == hipSuccess) e = launch(true, d_counts, d_off, d_out);
      if (e == hipSuccess) e = hipMemcpy(spmat, d_out, (size_t)total * 3 * sizeof(double), hipMemcpyDeviceToHost);
    }
  }
  if (e != hipSuccess) rc = hip_fail(e, "matrix generation");
  if (d_counts) (void)hipFree(d_counts);
  if (d_off) (void)hipFree(d_off);
  if (d_out) (void)hipFree(d_out);
  delete h;
  return rc != SGDCT_OK ? rc : (int)total;
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

sgdct_projector* sgdct_sf_init_offsets(int nx, int ny, int np, int nu, double dx, double dy, double du, double lsd,
                                       double lso, const double* angles, const double* dz, const double* drho) {
  (void)dz;  // out-of-plane source offset: no effect on a 2-D fan (z = 0 slice)
  return create(BASIS_SF, nx, ny, np, nu, dx, dy, du, lsd, lso, angles, drho);
}
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

// layers/A_B0.cu:449-492, A_B1.cu:545-593.
int pixel_matrix_generation(int np, int nu, int nx, int ny, double dx, double dy, double du, double lsd, double lso,
                            double* spmat, int buffer_size) {
  return matrix_generation(BASIS_B0, np, nu, nx, ny, dx, dy, du, lsd, lso, spmat, buffer_size);
}

int bilinear_matrix_generation(int np, int nu, int nx, int ny, double dx, double dy, double du, double lsd,
                               double lso, double* spmat, int buffer_size) {
  return matrix_generation(BASIS_B1, np, nu, nx, ny, dx, dy, du, lsd, lso, spmat, buffer_size);
}

// python/matrix_gen/A_H0_project.py and A_Hl_project.py bind these (the
// reference's own exporters, A_H0.cu:510-553 and A_Hl.cu:521-565, are
// commented out): nl rows x nc columns of lattice constant dt, np views over
// [0, 2 pi (np-1)/np], COO triplets (p*nu + u, il*nc + ic, conv).
int h0_matrix_generation(int np, int nu, int nl, int nc, double du, double dt, double lsd, double lso,
                         double* spmat, int buffer_size) {
  if (!(dt > 0)) return fail(SGDCT_EINVAL, "dt must be positive");
  return matrix_generation(BASIS_H0, np, nu, nc, nl, dt * std::sqrt(3.0), 1.5 * dt, du, lsd, lso, spmat,
                           buffer_size, dt);
}

int hl_matrix_generation(int np, int nu, int nl, int nc, double du, double dt, double lsd, double lso,
                         double* spmat, int buffer_size) {
  if (!(dt > 0)) return fail(SGDCT_EINVAL, "dt must be positive");
  return matrix_generation(BASIS_HL, np, nu, nc, nl, dt * std::sqrt(3.0), 0.5 * dt, du, lsd, lso, spmat,
                           buffer_size, dt);
}

int hex_matrix_generation(int np, int nu, int nl, int nc, double du, double dt, double lsd, double lso,
                          double* spmat, int buffer_size) {
  return hl_matrix_generation(np, nu, nl, nc, du, dt, lsd, lso, spmat, buffer_size);
}

// layers/A_B1_3D.hh:24-26 / A_B1_3D.cu:701-770.
sgdct_projector* b1_3d_init(int nx, int ny, int nz, int np, int nu, int nv, double dx, double dy, double dz,
                            double du, double dv, double lsd, double lso, const double* angles) {
  if (nz <= 0 || nv <= 0 || !(dz > 0) || !(dv > 0)) {
    fail(SGDCT_EINVAL, "invalid cone-beam geometry arguments");
    return nullptr;
  }
  sgdct_projector* h = create(BASIS_B1, nx, ny, np, nu, dx, dy, du, lsd, lso, angles, nullptr);
  if (!h) return nullptr;
  release_part(h);  // the cone-beam kernels do not split their launches
  h->basis = BASIS_B1_3D;
  h->nz = nz;
  h->nv = nv;
  h->dz = dz;
  h->dv = dv;
  return h;
}

unsigned long long b1_3d_forward_projection(double* b, double* x, sgdct_projector* h) {
  return (unsigned long long)cone_host_forward(b, x, h, BASIS_B1_3D);
}

int b1_3d_backward_projection(double* b, double* x, sgdct_projector* h) {
  return cone_host_backward(b, x, h, BASIS_B1_3D);
}

// layers/SF_projector.hh / SF_projector.cu:359-400 (class SF, no C entry
// points in the reference); trap_v selects SF-TT, else SF-TR.
sgdct_projector* sf3d_init(int nx, int ny, int nz, int np, int nu, int nv, double dx, double dy, double dz, double du,
                           double dv, double lsd, double lso, const double* angles, int trap_v) {
  if (nz <= 0 || nv <= 0 || !(dz > 0) || !(dv > 0)) {
    fail(SGDCT_EINVAL, "invalid cone-beam geometry arguments");
    return nullptr;
  }
  sgdct_projector* h = create(BASIS_SF, nx, ny, np, nu, dx, dy, du, lsd, lso, angles, nullptr, 0.0, 0.5);
  if (!h) return nullptr;
  release_part(h);
  h->basis = BASIS_SF_3D;
  h->cone_model = trap_v ? 2 : 1;
  h->nz = nz;
  h->nv = nv;
  h->dz = dz;
  h->dv = dv;
  return h;
}

int sf3d_forward_projection(double* b, double* x, sgdct_projector* h) {
  return (int)cone_host_forward(b, x, h, BASIS_SF_3D);
}

int sf3d_backward_projection(double* b, double* x, sgdct_projector* h) {
  return cone_host_backward(b, x, h, BASIS_SF_3D);
}

int sgdct_shape3(const sgdct_projector* h, int* out) {
  if (!h || !out) return fail(SGDCT_EINVAL, "null argument");
  out[0] = h->nx; out[1] = h->ny; out[2] = h->nz; out[3] = h->np; out[4] = h->nu; out[5] = h->nv;
  return h->basis;
}

void sgdct_destroy(sgdct_projector* h) { delete h; }

sgdct_batch* sgdct_batch_create(sgdct_projector* const* slices, int nslice) {
  if (!slices || nslice <= 0) {
    fail(SGDCT_EINVAL, "empty slice list");
    return nullptr;
  }
  auto* b = new (std::nothrow) sgdct_batch();
  if (!b) {
    fail(SGDCT_ENOMEM, "host allocation failed");
    return nullptr;
  }
  b->n = nslice;
  b->img_off.assign(nslice + 1, 0);
  b->sino_off.assign(nslice + 1, 0);
  for (int i = 0; i < nslice; ++i) {
    const sgdct_projector* h = slices[i];
    if (!h || (h->basis != BASIS_B0 && h->basis != BASIS_B1 && h->basis != BASIS_SF)) {
      fail(SGDCT_EUNSUP, "batched slices must be B0, B1 or SF projectors");
      delete b;
      return nullptr;
    }
    if (i > 0 && h->device != slices[0]->device) {
      fail(SGDCT_EINVAL, "batched slices must live on one device");
      delete b;
      return nullptr;
    }
    b->proj.push_back(slices[i]);
    SliceDesc d;
    d.g = h->args;
    d.basis = h->basis;
    d.lds = h->bwd_lds ? 1 : 0;
    d.img_off = b->img_off[i];
    d.sino_off = b->sino_off[i];
    b->desc.push_back(d);
    b->img_off[i + 1] = b->img_off[i] + (long long)h->nx * h->ny;
    b->sino_off[i + 1] = b->sino_off[i] + (long long)h->np * h->nu;
    const double w = basis_weight(h->basis);
    b->fwd_items.push_back((long long)h->np * h->args.nchunk);
    b->bwd_items.push_back((long long)((h->nx + kTile - 1) / kTile) * ((h->ny + kTile - 1) / kTile));
    b->fwd_cost.push_back(w * std::max(h->nx, h->ny));
    b->bwd_cost.push_back(w * h->np);
  }
  b->device = slices[0]->device;
  hipError_t e = hipSetDevice(b->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device);
  if (e == hipSuccess) e = hipMalloc(&b->d_slices, sizeof(SliceDesc) * nslice);
  if (e == hipSuccess) e = hipMemcpy(b->d_slices, b->desc.data(), sizeof(SliceDesc) * nslice, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&b->d_order, sizeof(int) * 2 * nslice);
  if (e == hipSuccess) e = hipMalloc(&b->d_prefix, sizeof(long long) * 2 * (nslice + 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_counters, sizeof(unsigned int) * 2);
  if (e != hipSuccess) {
    hip_fail(e, "batch allocation");
    delete b;
    return nullptr;
  }
  if (batch_upload_plan(b) != SGDCT_OK) {
    delete b;
    return nullptr;
  }
  if (b->plan.nfwd >= (1LL << 32) || b->plan.nbwd >= (1LL << 32)) {
    fail(SGDCT_EUNSUP, "batch has more than 2^32 work items");
    delete b;
    return nullptr;
  }
  return b;
}

void sgdct_batch_destroy(sgdct_batch* b) { delete b; }

int sgdct_batch_layout(const sgdct_batch* b, long long* img_off, long long* sino_off) {
  if (!b || !img_off || !sino_off) return fail(SGDCT_EINVAL, "null argument");
  for (int i = 0; i <= b->n; ++i) {
    img_off[i] = b->img_off[i];
    sino_off[i] = b->sino_off[i];
  }
  return b->n;
}

int sgdct_batch_config(sgdct_batch* b, int fwd_blocks_per_cu, int bwd_blocks_per_cu, int order) {
  if (!b || fwd_blocks_per_cu < 0 || fwd_blocks_per_cu > 32 || bwd_blocks_per_cu < 0 || bwd_blocks_per_cu > 32 ||
      (order != 0 && order != 1))
    return fail(SGDCT_EINVAL, "bad batch configuration");
  b->drop_graphs();
  b->fwd_bpc = fwd_blocks_per_cu;
  b->bwd_bpc = bwd_blocks_per_cu;
  if (order != b->order) {
    b->order = order;
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return hip_fail(e, "set device");
    return batch_upload_plan(b);
  }
  return SGDCT_OK;
}

int sgdct_batch_forward(sgdct_batch* b, const float* img, float* sino, void* stream) {
  if (!b || !img || !sino) return fail(SGDCT_EINVAL, "null argument");
  return batch_run(b, 0, img, sino, (hipStream_t)stream);
}

int sgdct_batch_backward(sgdct_batch* b, const float* sino, float* img, void* stream) {
  if (!b || !img || !sino) return fail(SGDCT_EINVAL, "null argument");
  return batch_run(b, 1, sino, img, (hipStream_t)stream);
}

int sgdct_batch_graph_launch(sgdct_batch* b, int direction, const float* in, float* out, void* stream) {
  if (!b || !in || !out || (direction != 0 && direction != 1)) return fail(SGDCT_EINVAL, "bad graph arguments");
  hipStream_t s = (hipStream_t)stream;
  if (!s) return fail(SGDCT_EINVAL, "graph capture needs a non-default stream");
  sgdct_batch::Graph& g = b->graph[direction];
  if (!g.exec || g.in != in || g.out != out) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = sgdct_batch::Graph{};
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "begin capture");
    int rc = batch_run(b, direction, in, out, s);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    if (rc != SGDCT_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "end capture");
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      g = sgdct_batch::Graph{};
      return hip_fail(e, "graph instantiate");
    }
    g.in = in;
    g.out = out;
  }
  hipError_t e = hipGraphLaunch(g.exec, s);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "graph launch");
}

int sgdct_split_config(sgdct_projector* h, int fwd_split, int bwd_split) {
  if (!h || fwd_split < 0 || fwd_split > kMaxSplit || bwd_split < 0 || bwd_split > kMaxSplit)
    return fail(SGDCT_EINVAL, "split_config: splits must be 0 (automatic) or 1..32");
  if (!splits(h->basis)) return fail(SGDCT_EINVAL, "split_config: fan-beam handles only");
  const int fs = h->fwd_split, bs = h->bwd_split;
  h->fwd_split = fwd_split;
  h->bwd_split = bwd_split;
  if (int rc = ensure_part(h, part_images(h) * part_need(h))) {
    h->fwd_split = fs;
    h->bwd_split = bs;
    return rc;
  }
  return SGDCT_OK;
}

int sgdct_forward_dev(sgdct_projector* h, const float* img, float* sino, int batch, void* stream) {
  if (!h || !img || !sino || batch <= 0) return fail(SGDCT_EINVAL, "bad forward_dev arguments");
  return run_forward(h, img, sino, batch, (hipStream_t)stream);
}

int sgdct_backward_dev(sgdct_projector* h, const float* sino, float* img, int batch, void* stream) {
  if (!h || !img || !sino || batch <= 0) return fail(SGDCT_EINVAL, "bad backward_dev arguments");
  return run_backward(h, sino, img, batch, (hipStream_t)stream);
}

int sgdct_project_dev_noalloc(sgdct_projector* h, int direction, const float* in, float* out, int batch,
                              void* stream) {
  if (!h || !in || !out || batch <= 0 || (direction != 0 && direction != 1))
    return fail(SGDCT_EINVAL, "bad project_dev_noalloc arguments");
  if (is_cone(h->basis))
    return fail(SGDCT_EUNSUP, "project_dev_noalloc: cone-beam handles stage their views in per-handle "
                              "workspace; use sgdct_forward_dev / sgdct_backward_dev");
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "get device");
  if (dev != h->device) return fail(SGDCT_EINVAL, "project_dev_noalloc: the handle's device is not current");
  double* part = nullptr;
  if (int rc = scratch_on(dev, &part)) return rc;
  // The handle's split, lowered (when it was chosen automatically) to the
  // largest whose partial sums of one image fit in the scratch array; a split
  // fixed by sgdct_split_config that does not fit is an error.
  int n = direction == 0 ? fwd_segments(h) : bwd_groups(h);
  const bool fixed = direction == 0 ? h->fwd_split != 0 : h->bwd_split != 0;
  const size_t unit = direction == 0 ? (size_t)h->np * h->nu : (size_t)h->nx * h->ny;
  if (n > 1 && (size_t)n * unit > kScratchDoubles) {
    if (fixed)
      return fail(SGDCT_EINVAL, "project_dev_noalloc: one image at the fixed split " + std::to_string(n) + " needs " +
                                    std::to_string((size_t)n * unit * sizeof(double)) +
                                    " bytes of partial sums, more than the " + std::to_string(kScratchBytes) +
                                    "-byte scratch array (lower it with sgdct_split_config)");
    n = (int)std::max<size_t>(1, kScratchDoubles / unit);
  }
  hipStream_t s = (hipStream_t)stream;
  return direction == 0 ? fan_forward(h, in, out, batch, n, part, kScratchDoubles, s)
                        : fan_backward(h, in, out, batch, n, part, kScratchDoubles, s);
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

// ---- warm-start network kernels (device pointers, asynchronous) ----------
namespace {
int nn_ret(hipError_t e, const char* what) { return e == hipSuccess ? SGDCT_OK : hip_fail(e, what); }
bool dims_ok(int B, int H, int W) { return B > 0 && H > 0 && W > 0; }
}  // namespace

int sgdct_nn_conv3x3(const void* in, const void* w, const float* bias, void* out, int B, int H, int W, int cin,
                     int cout, int relu, int out_f32, void* stream) {
  return sgdct_nn_conv3x3_ex(in, nullptr, 0, w, bias, out, nullptr, B, H, W, cin, cout, relu, out_f32, stream);
}

namespace {
// A two-source input (in at half resolution, c0 channels; in2 the skip
// tensor): c0 a positive multiple of 8 below cin, H and W even.
bool concat_ok(const void* in2, int c0, int H, int W, int cin) {
  return !in2 || (c0 > 0 && c0 < cin && c0 % 8 == 0 && H % 2 == 0 && W % 2 == 0);
}
}  // namespace

int sgdct_nn_conv3x3_ex(const void* in, const void* in2, int c0, const void* w, const float* bias, void* out,
                        const void* mask, int B, int H, int W, int cin, int cout, int relu, int out_f32,
                        void* stream) {
  if (!in || !w || !out || !dims_ok(B, H, W) || cin <= 0 || cin % 8 || cout <= 0)
    return fail(SGDCT_EINVAL, "bad conv3x3 arguments (cin must be a multiple of 8)");
  if (!concat_ok(in2, c0, H, W, cin))
    return fail(SGDCT_EINVAL, "bad conv3x3 concatenation (0 < c0 < cin, c0 % 8 == 0, even H and W)");
  return nn_ret(nn_conv3x3(in, in2, c0, w, bias, out, mask, B, H, W, cin, cout, relu, out_f32, (hipStream_t)stream),
                "conv3x3");
}

long long sgdct_nn_wgrad_workspace(int B, int H, int W, int cin, int cout) {
  if (!dims_ok(B, H, W) || cin <= 0 || cout <= 0) return fail(SGDCT_EINVAL, "bad wgrad shape");
  return nn_wgrad_splits((long long)B * H * W) * cout * (9LL * cin + 1);
}

int sgdct_nn_wgrad(const void* x, const void* dy, float* dw, float* workspace, long long workspace_floats, int B,
                   int H, int W, int cin, int cout, void* stream) {
  return sgdct_nn_wgrad_ex(x, nullptr, 0, dy, dw, nullptr, workspace, workspace_floats, B, H, W, cin, cout, stream);
}

int sgdct_nn_wgrad_ex(const void* x, const void* x2, int c0, const void* dy, float* dw, float* db, float* workspace,
                      long long workspace_floats, int B, int H, int W, int cin, int cout, void* stream) {
  if (!x || !dy || !dw || !workspace || !dims_ok(B, H, W) || cin <= 0 || cin % 8 || cout <= 0 || cout % 8)
    return fail(SGDCT_EINVAL, "bad wgrad arguments (cin and cout must be multiples of 8)");
  if (!concat_ok(x2, c0, H, W, cin))
    return fail(SGDCT_EINVAL, "bad wgrad concatenation (0 < c0 < cin, c0 % 8 == 0, even H and W)");
  if (workspace_floats < sgdct_nn_wgrad_workspace(B, H, W, cin, cout))
    return fail(SGDCT_EINVAL, "wgrad workspace too small");
  return nn_ret(nn_wgrad(x, x2, c0, dy, dw, db, workspace, B, H, W, cin, cout, (hipStream_t)stream), "wgrad");
}

long long sgdct_nn_colsum_workspace(long long M, int C) {
  if (M <= 0 || C <= 0) return fail(SGDCT_EINVAL, "bad colsum shape");
  return (M + 4095) / 4096 * (long long)C;
}

int sgdct_nn_colsum(const void* a, float* out, float* workspace, long long workspace_floats, long long M, int C,
                    void* stream) {
  if (!a || !out || !workspace || M <= 0 || C <= 0) return fail(SGDCT_EINVAL, "bad colsum arguments");
  if (workspace_floats < sgdct_nn_colsum_workspace(M, C)) return fail(SGDCT_EINVAL, "colsum workspace too small");
  return nn_ret(nn_colsum_bf16(a, out, workspace, M, C, (hipStream_t)stream), "colsum");
}

int sgdct_nn_pad_input(const float* x, void* out8, long long npix, void* stream) {
  if (!x || !out8 || npix <= 0) return fail(SGDCT_EINVAL, "bad pad_input arguments");
  return nn_ret(nn_pad_input(x, out8, npix, (hipStream_t)stream), "pad_input");
}

int sgdct_nn_pool2(const void* src, void* dst, int B, int H, int W, int C, void* stream) {
  if (!src || !dst || !dims_ok(B, H, W) || C <= 0) return fail(SGDCT_EINVAL, "bad pool2 arguments");
  return nn_ret(nn_pool2(src, dst, B, H, W, C, (hipStream_t)stream), "pool2");
}

int sgdct_nn_pool2_back(const float* dv, float* gsrc, void* out, const void* mask, int B, int H, int W, int C,
                        void* stream) {
  if (!dv || !gsrc || !dims_ok(B, H, W) || C <= 0 || (!out) != (!mask))
    return fail(SGDCT_EINVAL, "bad pool2_back arguments");
  return nn_ret(nn_pool2_back(dv, gsrc, out, mask, B, H, W, C, (hipStream_t)stream), "pool2_back");
}

int sgdct_nn_upcat(const void* low, const void* skip, void* dst, int B, int H, int W, int C0, int C1, void* stream) {
  if (!low || !skip || !dst || !dims_ok(B, H, W) || H % 2 || W % 2 || C0 <= 0 || C1 <= 0)
    return fail(SGDCT_EINVAL, "bad upcat arguments");
  return nn_ret(nn_upcat(low, skip, dst, B, H, W, C0, C1, (hipStream_t)stream), "upcat");
}

int sgdct_nn_upcat_back(const float* dv, float* glow, void* glow_out, const void* glow_mask, float* gskip,
                        int skip_accumulate, int B, int H, int W, int C0, int C1, void* stream) {
  if (!dv || !(glow_out ? glow_mask != nullptr : glow != nullptr) || !gskip || !dims_ok(B, H, W) || H % 2 || W % 2 ||
      C0 <= 0 || C1 <= 0)
    return fail(SGDCT_EINVAL, "bad upcat_back arguments");
  return nn_ret(nn_upcat_back(dv, glow, glow_out, glow_mask, gskip, skip_accumulate, B, H, W, C0, C1,
                              (hipStream_t)stream),
                "upcat_back");
}

int sgdct_nn_head(const void* d0, const float* w, const float* b, const float* x, float* y, long long M, int C,
                  void* stream) {
  if (!d0 || !w || !b || !x || !y || M <= 0 || C <= 0) return fail(SGDCT_EINVAL, "bad head arguments");
  return nn_ret(nn_head(d0, w, b, x, y, M, C, (hipStream_t)stream), "head");
}

long long sgdct_nn_head_workspace(long long M, int C) {
  if (M <= 0 || C <= 0) return fail(SGDCT_EINVAL, "bad head shape");
  return (M + 4095) / 4096 * (long long)(C + 1);
}

int sgdct_nn_head_back(const float* y, const float* target, const float* w, const void* d0, float* dy, float* gd0,
                       void* gd0_out, const void* mask, float* gw, float* workspace, long long workspace_floats,
                       long long M, int C, void* stream) {
  if (!y || !target || !w || !d0 || !dy || !(gd0_out ? mask != nullptr : gd0 != nullptr) || !gw || !workspace ||
      M <= 0 || C <= 0 || C > 255)
    return fail(SGDCT_EINVAL, "bad head_back arguments");
  if (workspace_floats < sgdct_nn_head_workspace(M, C)) return fail(SGDCT_EINVAL, "head workspace too small");
  return nn_ret(nn_head_back(y, target, w, d0, dy, gd0, gd0_out, mask, gw, workspace, M, C, (hipStream_t)stream),
                "head_back");
}

int sgdct_nn_relu_mask(const float* g, const void* act, void* out, long long n, void* stream) {
  if (!g || !act || !out || n <= 0) return fail(SGDCT_EINVAL, "bad relu_mask arguments");
  return nn_ret(nn_relu_mask(g, act, out, n, (hipStream_t)stream), "relu_mask");
}

int sgdct_nn_flip_weights(const void* w, void* wt, int cin, int cout, void* stream) {
  if (!w || !wt || cin <= 0 || cout <= 0) return fail(SGDCT_EINVAL, "bad flip_weights arguments");
  return nn_ret(nn_flip_weights(w, wt, cin, cout, (hipStream_t)stream), "flip_weights");
}

int sgdct_nn_adam(float* w32, void* w16, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                  float eps, int step, void* stream) {
  if (!w32 || !g || !m || !v || n <= 0 || step <= 0) return fail(SGDCT_EINVAL, "bad adam arguments");
  return nn_ret(nn_adam(w32, w16, g, m, v, n, lr, b1, b2, eps, step, (hipStream_t)stream), "adam");
}

int sgdct_nn_fill(float* p, long long n, float value, void* stream) {
  if (!p || n <= 0) return fail(SGDCT_EINVAL, "bad fill arguments");
  return nn_ret(nn_fill(p, n, value, (hipStream_t)stream), "fill");
}

int sgdct_nn_to_bf16(const float* src, void* dst, long long n, void* stream) {
  if (!src || !dst || n <= 0) return fail(SGDCT_EINVAL, "bad to_bf16 arguments");
  return nn_ret(nn_to_bf16(src, dst, n, (hipStream_t)stream), "to_bf16");
}

int sgdct_axpy_batch(float* x, const float* g, const float* a, long long n, long long per, void* stream) {
  if (!x || !g || !a || n <= 0 || per <= 0 || n % per) return fail(SGDCT_EINVAL, "bad axpy_batch arguments");
  hipError_t e = launch_axpy_batch(x, g, a, (size_t)n, (size_t)per, (hipStream_t)stream);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "axpy_batch launch");
}

int sgdct_tv_descent(float* x, float* g, double* work, const double* step, int nx, int ny, int batch, float eps,
                     int nsteps, void* stream) {
  if (!x || !g || !work || !step || nx <= 0 || ny <= 0 || batch <= 0 || !(eps > 0) || nsteps < 0)
    return fail(SGDCT_EINVAL, "bad tv_descent arguments");
  if (nsteps == 0) return SGDCT_OK;
  hipError_t e = launch_tv_descent(x, g, work, step, nx, ny, batch, eps, nsteps, (hipStream_t)stream);
  return e == hipSuccess ? SGDCT_OK : hip_fail(e, "tv_descent launch");
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
