/*
 * sgdct_oracle.c -- CPU restatement of the reference's fan-beam footprint
 * projectors (B0, B1, SF, H0, Hl) and an FP64 strip-integral oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only to check results
 * or as the timed CPU baseline.  The product (sgd_amd/) never links or calls it.
 *
 * The file is compiled twice (see oracle/Makefile):
 *   -DORACLE_FP=float   "reference mode": the footprint arithmetic is FP32
 *                       exactly where layers/A_*.cu compute in float, with the
 *                       same expression forms and evaluation order, and FP64
 *                       accumulation where the reference accumulates in double.
 *   -DORACLE_FP=double  "exact mode": the same formulas evaluated in FP64.
 * The closed-form polynomials below are kept term-for-term equal to the
 * reference's (layers/spline_conv.hh) on purpose: reference mode is meant to
 * reproduce the reference's FP32 rounding, which depends on expression form.
 *
 * Image layout: img[iy*nx + ix]; sinogram layout: sino[p*nu + u].
 *
 * Pins (see tests/test_oracle_*.py): closed-form <-> truncated-power
 * identities, unit integrals and symmetry of every 1-D footprint, brute-force
 * exact strip integrals (or_strip_exact) on small inputs, adjointness,
 * 90-degree rotation <-> view shift, per-view mass.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORACLE_FP
#define ORACLE_FP float
#endif
typedef ORACLE_FP fpt;

#define OR_PI 3.14159265358979323846
#define EPS7 ((fpt)1e-7)

static inline fpt fp_abs(fpt v) { return v < 0 ? -v : v; }
static inline fpt fp_max(fpt a, fpt b) { return a > b ? a : b; }
static inline fpt fp_min(fpt a, fpt b) { return a < b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int fp_floor_int(fpt v) { return (int)floor((double)v); }

/* ------------------------------------------------------------------------- */
/* Geometry: restates utils/geo.cu:232-269 (source, detector centre and      */
/* detector step per view, 45-degree snapping) and the projective map of     */
/* utils/geo.cu:58-161 in closed form.                                       */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, np, nu;
    double dx, dy, du, lsd, lso;
    double *src; /* np*2 */
    double *dtv; /* np*2 */
    double *puv; /* np*2 */
    double *pm;  /* np*12, projective map (x, y, z, 1) -> u, rows 0..2 */
    fpt *pmi;    /* np*12, pm pre-multiplied for integer pixel indices */
} ogeo;

/* Per-view source/detector placement (utils/geo.cu:240-262). */
static void view_placement(double angle, double lsd, double lso, double du, double drho,
                           double *src, double *dtv, double *puv) {
    double beta = angle;
    double s = sin(beta), c = cos(beta);
    /* snap views within 1e-7 of the 45-degree diagonals (utils/geo.cu:254-257) */
    if (fabs(s - c) < 1e-7)
        beta = (s > 0) ? OR_PI / 4 : -OR_PI * 3 / 4;
    else if (fabs(s + c) < 1e-7)
        beta = (s > 0) ? OR_PI * 3 / 4 : -OR_PI / 4;
    s = sin(beta);
    c = cos(beta);
    double sx = -lso + drho;
    src[0] = sx * c;
    src[1] = sx * s;
    dtv[0] = (lsd - lso) * c;
    dtv[1] = (lsd - lso) * s;
    puv[0] = -du * s;
    puv[1] = du * c;
}

/* Closed-form projective map u(X) = nu/2 + (L/du) ((X-src).t)/((X-src).e)
 * with e the unit source->detector axis, t = puv/du, L the source-detector
 * distance; rows scaled so that pm[11] == 1 (utils/geo.cu:135).  The
 * reference builds the same map by matrix products and a 3x3 inverse, with a
 * special identity fallback on the 3pi/4, 7pi/4 diagonals that evaluates to
 * the same u; the closed form needs no fallback. */
/* Index-space map terms as the reference forms them (utils/geo.cu:141-159):
 * the grid step d and the centre offset (n - 1) d / 2 are float there, the
 * offset a float product.  Reference mode keeps both roundings; exact mode
 * evaluates them in FP64 like every other term (with d = 0.15 mm and
 * n = 2048 the float offset alone moves the image by 6e-6 mm). */
static double grid_step(double d) { return (double)(fpt)d; }
static double centre_offset(int n, double d) { return (double)((fpt)(n - 1) * (fpt)d / (fpt)2.0f); }

static void view_pm(const double *src, const double *dtv, const double *puv, double du, int nu,
                    int nx, int ny, double dx, double dy, double *pm, fpt *pmi) {
    double ex = dtv[0] - src[0], ey = dtv[1] - src[1];
    double L = sqrt(ex * ex + ey * ey);
    ex /= L;
    ey /= L;
    double tx = puv[0] / du, ty = puv[1] / du;
    double uc = nu / 2.0, k = L / du;
    double d0 = -(src[0] * ex + src[1] * ey);   /* constant term of row 2 */
    double n0 = -(src[0] * tx + src[1] * ty);
    double row0[4] = {uc * ex + k * tx, uc * ey + k * ty, 0.0, uc * d0 + k * n0};
    double row1[4] = {0.5 * ex, 0.5 * ey, k, 0.5 * d0};
    double row2[4] = {ex, ey, 0.0, d0};
    for (int j = 0; j < 4; ++j) {
        pm[j] = row0[j] / d0;
        pm[4 + j] = row1[j] / d0;
        pm[8 + j] = row2[j] / d0;
    }
    /* index-space version (utils/geo.cu:151-160) */
    const double sx = grid_step(dx), sy = grid_step(dy);
    const double fox = centre_offset(nx, dx), foy = centre_offset(ny, dy);
    for (int r = 0; r < 3; ++r) {
        pmi[4 * r + 0] = (fpt)(sx * pm[4 * r + 0]);
        pmi[4 * r + 1] = (fpt)(sy * pm[4 * r + 1]);
        pmi[4 * r + 2] = (fpt)(1.0 * pm[4 * r + 2]);
        pmi[4 * r + 3] = (fpt)(pm[4 * r + 3] - fox * pm[4 * r + 0] - foy * pm[4 * r + 1]);
    }
}

void *or_geo_create(int nx, int ny, int np, int nu, double dx, double dy, double du, double lsd,
                    double lso, const double *angles, const double *drho) {
    ogeo *g = (ogeo *)calloc(1, sizeof(ogeo));
    g->nx = nx; g->ny = ny; g->np = np; g->nu = nu;
    g->dx = dx; g->dy = dy; g->du = du; g->lsd = lsd; g->lso = lso;
    g->src = (double *)malloc(sizeof(double) * 2 * np);
    g->dtv = (double *)malloc(sizeof(double) * 2 * np);
    g->puv = (double *)malloc(sizeof(double) * 2 * np);
    g->pm = (double *)malloc(sizeof(double) * 12 * np);
    g->pmi = (fpt *)malloc(sizeof(fpt) * 12 * np);
    for (int p = 0; p < np; ++p) {
        view_placement(angles[p], lsd, lso, du, drho ? drho[p] : 0.0, g->src + 2 * p,
                       g->dtv + 2 * p, g->puv + 2 * p);
        view_pm(g->src + 2 * p, g->dtv + 2 * p, g->puv + 2 * p, du, nu, nx, ny, dx, dy,
                g->pm + 12 * p, g->pmi + 12 * p);
    }
    return g;
}

void or_geo_destroy(void *h) {
    ogeo *g = (ogeo *)h;
    if (!g) return;
    free(g->src); free(g->dtv); free(g->puv); free(g->pm); free(g->pmi);
    free(g);
}

/* Exported for the geometry tests: continuous detector coordinate of (x, y). */
double or_geo_u(const void *h, int p, double x, double y) {
    const ogeo *g = (const ogeo *)h;
    const double *pm = g->pm + 12 * p;
    return (pm[0] * x + pm[1] * y + pm[3]) / (pm[8] * x + pm[9] * y + pm[11]);
}

void or_geo_view(const void *h, int p, double *src, double *dtv, double *puv) {
    const ogeo *g = (const ogeo *)h;
    src[0] = g->src[2 * p]; src[1] = g->src[2 * p + 1];
    dtv[0] = g->dtv[2 * p]; dtv[1] = g->dtv[2 * p + 1];
    puv[0] = g->puv[2 * p]; puv[1] = g->puv[2 * p + 1];
}

/* Bin range of a pixel from the projections of four corners at index offset
 * +-h (h = 1 for B1, 1/2 for B0 and SF), sorted, floored and clamped
 * (layers/A_B1.cu:100-151, layers/A_B0.cu:158-195).  us[] gets the sorted
 * corner projections (SF uses them as its trapezoid). */
static void corner_range(const fpt *pm, int ix, int iy, fpt h, int nu, fpt *us, int *minu,
                         int *maxu) {
    fpt sx1 = (fpt)ix - h, sx2 = (fpt)ix + h;
    fpt sy1 = (fpt)iy - h, sy2 = (fpt)iy + h;
    us[0] = (pm[0] * sx1 + pm[3] + pm[1] * sy1) / (pm[8] * sx1 + pm[11] + pm[9] * sy1);
    us[1] = (pm[0] * sx1 + pm[3] + pm[1] * sy2) / (pm[8] * sx1 + pm[11] + pm[9] * sy2);
    us[2] = (pm[0] * sx2 + pm[3] + pm[1] * sy1) / (pm[8] * sx2 + pm[11] + pm[9] * sy1);
    us[3] = (pm[0] * sx2 + pm[3] + pm[1] * sy2) / (pm[8] * sx2 + pm[11] + pm[9] * sy2);
    /* sorting network of layers/spline_conv.hh:24-39 */
#define OR_SORT2(a, b) do { if ((a) > (b)) { fpt t_ = (b); (b) = (a); (a) = t_; } } while (0)
    OR_SORT2(us[0], us[1]);
    OR_SORT2(us[2], us[3]);
    OR_SORT2(us[0], us[2]);
    OR_SORT2(us[1], us[3]);
    OR_SORT2(us[1], us[2]);
#undef OR_SORT2
    *minu = imin(imax(0, fp_floor_int(us[0])), nu - 1);
    *maxu = imin(imax(0, fp_floor_int(us[3])), nu - 1);
}

/* ------------------------------------------------------------------------- */
/* 1-D footprints (layers/spline_conv.hh)                                     */
/* ------------------------------------------------------------------------- */

/* univariate3dirboxspline_fast2, spline_conv.hh:239-242 (xi sorted desc.) */
static fpt box3_fast2(fpt x, const fpt *xi) {
    return fp_max((fpt)0.0f, fp_min((fpt)1.0f / xi[0], ((xi[0] + xi[1]) / (fpt)2.0f - x) / (xi[0] * xi[1])));
}

/* univariate3dirboxspline_fast3, spline_conv.hh:244-293: three unit-integral
 * boxes of widths xi[0] >= xi[1] >= xi[2] convolved, x = |offset from centre| */
static fpt box3_fast3(fpt x, const fpt *xi) {
    fpt h = (fpt)1.0f / xi[0];
    fpt p = xi[0] * xi[1] * xi[2];
    fpt K0 = (xi[0] - xi[1] - xi[2]) / (fpt)2.0f;
    fpt K1 = (xi[0] - xi[1] + xi[2]) / (fpt)2.0f;
    fpt K2 = (xi[0] + xi[1] - xi[2]) / (fpt)2.0f;
    fpt K3 = (xi[0] + xi[1] + xi[2]) / (fpt)2.0f;
    if (x >= K3) return (fpt)0.0f;
    if (x >= K2) return (fpt)0.5f * (K3 - x) * (K3 - x) / p;
    if (x >= K1) return ((xi[0] + xi[1]) / (fpt)2.0f - x) * xi[2] / p;
    if (x < -K0) return h - (K0 * K0 + x * x) / p;
    if (x < K0) return h;
    return h - (fpt)0.5f * (K0 - x) * (K0 - x) / p;
}

/* B0 footprint: layers/A_B0.cu:20-35 (linear_convolution_1d). */
static fpt b0_conv(fpt y, fpt v1, fpt v2, fpt v3) {
    fpt vs[3];
    y = fp_abs(y - (v1 + v2 + v3) / (fpt)2.0f);
    vs[0] = fp_abs(v1); vs[1] = fp_abs(v2); vs[2] = fp_abs(v3);
    /* sort3d, descending (spline_conv.hh:52-58) */
#define OR_SORT2D(a, b) do { if ((a) < (b)) { fpt t_ = (b); (b) = (a); (a) = t_; } } while (0)
    OR_SORT2D(vs[1], vs[2]);
    OR_SORT2D(vs[0], vs[1]);
    OR_SORT2D(vs[1], vs[2]);
#undef OR_SORT2D
    if (vs[2] > EPS7) return box3_fast3(y, vs);
    return box3_fast2(y, vs);
}

/* Antiderivative from 0 of tri_a * tri_b (unit-integral hats, half-widths
 * a >= b), branch a > 2b: spline_conv.hh:295-332. */
static fpt s2tri_ag2b(fpt a, fpt b, fpt x) {
    fpt a3 = a * a * a, ab2 = a * b * b, b3 = b * b * b;
    fpt k1 = (6 * a * a * b * b);
    fpt sum = (fpt)0.0f;
    if (x <= b) return (x * (12 * ab2 - 4 * b3 - 4 * b * x * x + x * x * x)) / (2 * k1);
    sum += ((12 * a - 7 * b) * b) / (12 * a * a);
    if (x <= a - b) return sum + (-2 * a * b + b * b + 2 * a * x - x * x) / (2 * a * a);
    sum += (fpt)0.5f - b / a;
    if (x <= a)
        return sum + ((a - b - x) * (a3 - 13 * b3 + 9 * b * b * x - 3 * b * x * x - x * x * x -
                                     3 * a * a * (b + x) - 3 * a * (b - x) * (3 * b + x))) / (4 * k1);
    sum += (13 * b * b) / (24 * a * a);
    if (x < a + b)
        return sum + ((x - a) * (a + 2 * b - x) * (a * a + 2 * a * b + 2 * b * b - 2 * (a + b) * x + x * x)) / (4 * k1);
    return sum + b * b / (24 * a * a);
}

/* Branch a <= 2b: spline_conv.hh:334-370. */
static fpt s2tri_al2b(fpt a, fpt b, fpt x) {
    fpt a3 = a * a * a, a2b = a * a * b, ab2 = a * b * b, b3 = b * b * b;
    fpt k1 = (6 * a * a * b * b);
    fpt k2 = a - b;
    fpt sum = (fpt)0.0f;
    if (x <= a - b) return (x * (12 * ab2 - 4 * b3 - 4 * b * x * x + x * x * x)) / (2 * k1);
    sum += ((a - b) * (a3 - 7 * a2b + 23 * ab2 - 9 * b3)) / (2 * k1);
    if (x <= b)
        return sum + (((x - a + b) * (a3 - 11 * a2b + 43 * ab2 - 17 * b3 + (5 * a - 13 * b) * k2 * x -
                                      (a + 7 * b) * x * x + 3 * x * x * x)) / (4 * k1));
    sum += (((2 * b - a) * (a3 - 6 * a2b + 24 * ab2 - 8 * b3)) / (4 * k1));
    if (x <= a)
        return sum + (((b - x) * (4 * a3 - 3 * b3 + b * b * x - 5 * b * x * x - x * x * x -
                                  6 * a * a * (3 * b + x) + 4 * a * (b * b + 4 * b * x + x * x))) / (4 * k1));
    sum += (((b - a) * (a3 - 7 * a2b + 5 * ab2 - 3 * b3)) / (4 * k1));
    if (x < a + b)
        return sum + ((x - a) * (a + 2 * b - x) * (a * a + 2 * a * b + 2 * b * b - 2 * (a + b) * x + x * x)) / (4 * k1);
    return sum + b * b / (24 * a * a);
}

/* fast_2tri_1box_spline, spline_conv.hh:372-390 */
static fpt tri2box(fpt a, fpt b, fpt x1, fpt x2) {
    fpt s1, s2;
    int side = ((x1 * x2) > 0);
    x1 = fp_abs(x1); x2 = fp_abs(x2);
    if (b <= a - b) { s1 = s2tri_ag2b(a, b, x1); s2 = s2tri_ag2b(a, b, x2); }
    else { s1 = s2tri_al2b(a, b, x1); s2 = s2tri_al2b(a, b, x2); }
    return side ? fp_abs(s1 - s2) : (s1 + s2);
}

/* fast_1tri_1box_spline, spline_conv.hh:393-405 */
static fpt tri1box(fpt a, fpt x1, fpt x2) {
    int side = ((x1 * x2) > 0);
    x1 = fp_abs(x1); x2 = fp_abs(x2);
    fpt s1 = (x1 >= a) ? (fpt)0.5f : (x1 / a - (x1 * x1) / (2 * a * a));
    fpt s2 = (x2 >= a) ? (fpt)0.5f : (x2 / a - (x2 * x2) / (2 * a * a));
    return side ? fp_abs(s1 - s2) : (s1 + s2);
}

/* B1 footprint: layers/A_B1.cu:48-59 (linear_convolution_1d). */
static fpt b1_conv(fpt x1, fpt x2, fpt a, fpt b) {
    if (b > EPS7) return tri2box(a, b, x1, x2) / fp_abs(x2 - x1);
    return tri1box(a, x1, x2) / fp_abs(x2 - x1);
}

/* exported primitives for the footprint tests */
double or_b0_conv(double y, double v1, double v2, double v3) {
    return (double)b0_conv((fpt)y, (fpt)v1, (fpt)v2, (fpt)v3);
}
double or_b1_conv(double x1, double x2, double a, double b) {
    return (double)b1_conv((fpt)x1, (fpt)x2, (fpt)a, (fpt)b);
}
double or_s2tri(double a, double b, double x) {
    return (double)((b <= a - b) ? s2tri_ag2b((fpt)a, (fpt)b, (fpt)x) : s2tri_al2b((fpt)a, (fpt)b, (fpt)x));
}
double or_box3(double y, double v1, double v2, double v3) {
    /* centred three-direction box spline value at |y| (no re-centring) */
    fpt vs[3] = {(fpt)fabs(v1), (fpt)fabs(v2), (fpt)fabs(v3)};
    fpt t;
    if (vs[0] < vs[1]) { t = vs[0]; vs[0] = vs[1]; vs[1] = t; }
    if (vs[1] < vs[2]) { t = vs[1]; vs[1] = vs[2]; vs[2] = t; }
    if (vs[0] < vs[1]) { t = vs[0]; vs[0] = vs[1]; vs[1] = t; }
    if (vs[2] > EPS7) return box3_fast3((fpt)fabs(y), vs);
    return box3_fast2((fpt)fabs(y), vs);
}

/* ------------------------------------------------------------------------- */
/* Per-(view, bin) tables                                                     */
/* ------------------------------------------------------------------------- */

/* B1 vx, vy: layers/A_B1.cu:429-478 (FP64, stored as fpt). */
static void b1_vectors(const ogeo *g, int p, fpt *vx, fpt *vy) {
    double cx = g->dtv[2 * p], cy = g->dtv[2 * p + 1];
    double ux = g->puv[2 * p], uy = g->puv[2 * p + 1];
    double px = g->src[2 * p], py = g->src[2 * p + 1];
    for (int u = 0; u < g->nu; ++u) {
        double sx = cx + ux * (u - 0.5 * g->nu + 0.5);
        double sy = cy + uy * (u - 0.5 * g->nu + 0.5);
        double th = atan2(fabs(px - sx), fabs(py - sy));
        vx[u] = (fpt)fabs(g->dx * cos(th));
        vy[u] = (fpt)fabs(g->dy * sin(th));
    }
}

/* B0 signed vx, vy and bin-edge offsets: layers/A_B0.cu:338-388 (FP64). */
static void b0_vectors(const ogeo *g, int p, double *tab /* 5*nu */) {
    int nu = g->nu;
    double cx = g->dtv[2 * p], cy = g->dtv[2 * p + 1];
    double ux = g->puv[2 * p], uy = g->puv[2 * p + 1];
    double px = g->src[2 * p], py = g->src[2 * p + 1];
    for (int u = 0; u < nu; ++u) {
        double sx = cx + ux * (u - 0.5 * nu + 0.5);
        double sy = cy + uy * (u - 0.5 * nu + 0.5);
        double ax = sx - ux * 0.5, ay = sy - uy * 0.5;
        double bx = sx + ux * 0.5, by = sy + uy * 0.5;
        double lsp = sqrt((px - sx) * (px - sx) + (py - sy) * (py - sy));
        double th = atan2(fabs(px - sx), fabs(py - sy));
        double vx = fabs(g->dx * cos(th)), vy = fabs(g->dy * sin(th));
        if (py < sy) vx = -vx;
        if (px < sx) vy = -vy;
        double a0s = fabs((sx - px) * (ay - py) - (ax - px) * (sy - py)) / lsp;
        double sb0 = fabs((sx - px) * (by - py) - (bx - px) * (sy - py)) / lsp;
        a0s /= ((ax - px) * (sx - px) + (ay - py) * (sy - py)) / (lsp * lsp);
        sb0 /= ((bx - px) * (sx - px) + (by - py) * (sy - py)) / (lsp * lsp);
        tab[u] = vx; tab[nu + u] = vy; tab[2 * nu + u] = a0s; tab[3 * nu + u] = sb0; tab[4 * nu + u] = lsp;
    }
}

/* ------------------------------------------------------------------------- */
/* Per-pixel, per-view footprint enumeration                                  */
/* Each basis enumerates (bin, weight) with weight such that                  */
/*   A: sino[p][u] += weight * img[pix]       A^T: img[pix] += weight * sino  */
/* and follows the reference's skip rules (conv > 1e-7, NaN guard).           */
/* ------------------------------------------------------------------------- */
typedef void (*emit_fn)(void *ctx, int u, double w);

/* B1 for one (pixel, view): layers/A_B1.cu:62-208 (forward) / :211-353. */
static void b1_pixel(const ogeo *g, int p, const fpt *vx, const fpt *vy, int ix, int iy, int pad,
                     emit_fn emit, void *ctx) {
    int nx = g->nx, ny = g->ny, nu = g->nu;
    fpt dx = (fpt)g->dx, dy = (fpt)g->dy;
    fpt px = (fpt)g->src[2 * p], py = (fpt)g->src[2 * p + 1];
    fpt pux = (fpt)g->puv[2 * p], puy = (fpt)g->puv[2 * p + 1];
    fpt dtx = (fpt)g->dtv[2 * p], dty = (fpt)g->dtv[2 * p + 1];
    fpt us[4];
    int minu, maxu;
    fpt ox = ((fpt)ix - (fpt)0.5f * (fpt)(nx - 1)) * dx;
    fpt oy = ((fpt)iy - (fpt)0.5f * (fpt)(ny - 1)) * dy;
    corner_range(g->pmi + 12 * p, ix, iy, (fpt)1.0f, nu, us, &minu, &maxu);
    minu = imax(0, minu - pad);
    maxu = imin(nu - 1, maxu + pad);
    fpt sx = dtx + pux * ((fpt)minu - (fpt)0.5f * (fpt)nu + (fpt)0.5f);
    fpt sy = dty + puy * ((fpt)minu - (fpt)0.5f * (fpt)nu + (fpt)0.5f);
    for (int u = minu; u <= maxu; ++u, sx += pux, sy += puy) {
        fpt ax = sx - pux * (fpt)0.5f, ay = sy - puy * (fpt)0.5f;
        fpt bx = sx + pux * (fpt)0.5f, by = sy + puy * (fpt)0.5f;
        fpt ra = ((ox - px) * (sx - px) + (oy - py) * (sy - py)) / ((ax - px) * (sx - px) + (ay - py) * (sy - py));
        fpt rb = ((ox - px) * (sx - px) + (oy - py) * (sy - py)) / ((bx - px) * (sx - px) + (by - py) * (sy - py));
        fpt ax1 = px + (ax - px) * ra, ay1 = py + (ay - py) * ra;
        fpt bx1 = px + (bx - px) * rb, by1 = py + (by - py) * rb;
        fpt x1 = (fpt)sqrt((double)((ax1 - ox) * (ax1 - ox) + (ay1 - oy) * (ay1 - oy)));
        fpt x2 = (fpt)sqrt((double)((bx1 - ox) * (bx1 - ox) + (by1 - oy) * (by1 - oy)));
        fpt sgn = ((ax1 - ox) * (bx1 - ox) + (ay1 - oy) * (by1 - oy) > 0) ? (fpt)1.0f : (fpt)-1.0f;
        fpt a = fp_max(vx[u], vy[u]), b = fp_min(vx[u], vy[u]);
        fpt conv = b1_conv(x1 * sgn, x2, a, b);
        if (conv == conv && conv > EPS7) emit(ctx, u, (double)conv * (double)dx * (double)dy);
    }
}

/* B0 for one (pixel, view): layers/A_B0.cu:123-229 / :231-336. */
static void b0_pixel(const ogeo *g, int p, const double *tab, int ix, int iy, int pad,
                     emit_fn emit, void *ctx) {
    int nx = g->nx, ny = g->ny, nu = g->nu;
    fpt dx = (fpt)g->dx, dy = (fpt)g->dy;
    fpt nx2 = (fpt)0.5f * (fpt)(nx - 1), ny2 = (fpt)0.5f * (fpt)(ny - 1);
    fpt px = (fpt)g->src[2 * p], py = (fpt)g->src[2 * p + 1];
    fpt pux = (fpt)g->puv[2 * p], puy = (fpt)g->puv[2 * p + 1];
    fpt dtx = (fpt)g->dtv[2 * p], dty = (fpt)g->dtv[2 * p + 1];
    fpt us[4];
    int minu, maxu;
    fpt ox = ((fpt)ix - nx2) * dx;
    fpt oy = ((fpt)iy - ny2) * dy;
    corner_range(g->pmi + 12 * p, ix, iy, (fpt)0.5f, nu, us, &minu, &maxu);
    minu = imax(0, minu - pad);
    maxu = imin(nu - 1, maxu + pad);
    fpt sx = dtx + pux * ((fpt)minu - (fpt)0.5f * (fpt)nu + (fpt)0.5f);
    fpt sy = dty + puy * ((fpt)minu - (fpt)0.5f * (fpt)nu + (fpt)0.5f);
    for (int u = minu; u <= maxu; ++u, sx += pux, sy += puy) {
        double vx = tab[u], vy = tab[nu + u], a0s = tab[2 * nu + u], sb0 = tab[3 * nu + u], lsp = tab[4 * nu + u];
        double r1 = (double)((ox - px) * (sx - px) + (oy - py) * (sy - py)) / (lsp * lsp);
        double vblur = (a0s + sb0) * (-r1);
        fpt zx = ox - (fpt)0.5f * dx - px, zy = oy + (fpt)0.5f * dy - py;
        double y = (double)((sx - px) * zy - zx * (sy - py)) / lsp;
        y += a0s * r1;
        fpt conv = b0_conv((fpt)(-y), (fpt)vx, (fpt)vy, (fpt)vblur);
        if (conv == conv && conv > EPS7) emit(ctx, u, (double)conv * (double)dx * (double)dy);
    }
}

/* SF trapezoid/bin overlap: layers/A_SF.cu:48-84. */
static fpt sf_gamma(fpt s1, fpt s2, const fpt *us) {
    fpt acc = (fpt)0.0f, tmp, b1, b2;
    b1 = fp_max(s1, us[0]); b2 = fp_min(s2, us[1]);
    tmp = (b2 > b1) ? ((fpt)1 / (2 * (us[1] - us[0]))) * ((b2 - us[0]) * (b2 - us[0]) - (b1 - us[0]) * (b1 - us[0])) : (fpt)0.0f;
    if (tmp == tmp) acc += tmp;
    b1 = fp_max(s1, us[1]); b2 = fp_min(s2, us[2]);
    tmp = (b2 > b1) ? (b2 - b1) : (fpt)0.0f;
    if (tmp == tmp) acc += tmp;
    b1 = fp_max(s1, us[2]); b2 = fp_min(s2, us[3]);
    tmp = (b2 > b1) ? ((fpt)1 / (2 * (us[3] - us[2]))) * ((b1 - us[3]) * (b1 - us[3]) - (b2 - us[3]) * (b2 - us[3])) : (fpt)0.0f;
    if (tmp == tmp) acc += tmp;
    return acc;
}

/* SF for one (pixel, view): layers/A_SF.cu:87-225 / :228-334. */
static void sf_pixel(const ogeo *g, int p, int ix, int iy, int pad, emit_fn emit, void *ctx) {
    int nx = g->nx, ny = g->ny, nu = g->nu;
    fpt dx = (fpt)g->dx, dy = (fpt)g->dy;
    fpt nx2 = (fpt)(nx - 1) / (fpt)2.0f, ny2 = (fpt)(ny - 1) / (fpt)2.0f;
    fpt px = (fpt)g->src[2 * p], py = (fpt)g->src[2 * p + 1];
    fpt us[4];
    int minu, maxu;
    corner_range(g->pmi + 12 * p, ix, iy, (fpt)0.5f, nu, us, &minu, &maxu);
    minu = imax(0, minu - pad);
    maxu = imin(nu - 1, maxu + pad);
    fpt th = (fpt)atan2((double)(dx * ((fpt)ix - nx2) - px), (double)(dy * ((fpt)iy - ny2) - py));
    fpt st = fp_abs((fpt)sin((double)th)), ct = fp_abs((fpt)cos((double)th));
    fpt weight = dx / fp_max(st, ct);
    for (int u = minu; u <= maxu; ++u) {
        fpt gm = sf_gamma((fpt)u, (fpt)u + (fpt)1.0f, us);
        double w = (double)weight * (double)gm;
        if (w == w && gm > EPS7) emit(ctx, u, w);
    }
}

/* ------------------------------------------------------------------------- */
/* Whole-image drivers (OpenMP: A over views, A^T over pixels).              */
/* ------------------------------------------------------------------------- */
typedef struct { double *row; double c; } fwd_ctx;
static void emit_fwd(void *ctx, int u, double w) {
    fwd_ctx *f = (fwd_ctx *)ctx;
    f->row[u] += w * f->c;
}
typedef struct { const double *row; double acc; } bwd_ctx;
static void emit_bwd(void *ctx, int u, double w) {
    bwd_ctx *b = (bwd_ctx *)ctx;
    b->acc += w * b->row[u];
}
typedef struct { int *u; double *w; int n, cap; } list_ctx;
static void emit_list(void *ctx, int u, double w) {
    list_ctx *l = (list_ctx *)ctx;
    if (l->n < l->cap) { l->u[l->n] = u; l->w[l->n] = w; }
    l->n++;
}

enum { OR_B0 = 0, OR_B1 = 1, OR_SF = 2 };

static void pixel_dispatch(int basis, const ogeo *g, int p, const void *tab, int ix, int iy, int pad,
                           emit_fn e, void *ctx) {
    if (basis == OR_B1) {
        const fpt *vx = (const fpt *)tab;
        b1_pixel(g, p, vx, vx + g->nu, ix, iy, pad, e, ctx);
    } else if (basis == OR_B0) {
        b0_pixel(g, p, (const double *)tab, ix, iy, pad, e, ctx);
    } else {
        sf_pixel(g, p, ix, iy, pad, e, ctx);
    }
}

static void *make_tab(int basis, const ogeo *g, int p) {
    if (basis == OR_B1) {
        fpt *t = (fpt *)malloc(sizeof(fpt) * 2 * g->nu);
        b1_vectors(g, p, t, t + g->nu);
        return t;
    }
    if (basis == OR_B0) {
        double *t = (double *)malloc(sizeof(double) * 5 * g->nu);
        b0_vectors(g, p, t);
        return t;
    }
    return NULL;
}

/* A: sino[p*nu+u] = sum over pixels; views [p0, p1) only (others untouched). */
void or_forward(int basis, const void *h, const double *img, double *sino, int p0, int p1, int pad) {
    const ogeo *g = (const ogeo *)h;
    int nx = g->nx, ny = g->ny, nu = g->nu;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = p0; p < p1; ++p) {
        void *tab = make_tab(basis, g, p);
        double *row = sino + (size_t)p * nu;
        memset(row, 0, sizeof(double) * nu);
        fwd_ctx f;
        f.row = row;
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                f.c = img[(size_t)iy * nx + ix];
                pixel_dispatch(basis, g, p, tab, ix, iy, pad, emit_fwd, &f);
            }
        free(tab);
    }
}

/* A^T over views [p0, p1): img[pix] = sum_p sum_u w * sino[p][u]. */
void or_backward(int basis, const void *h, const double *sino, double *img, int p0, int p1, int pad) {
    const ogeo *g = (const ogeo *)h;
    int nx = g->nx, ny = g->ny, nu = g->nu;
    int np = p1 - p0;
    void **tabs = (void **)malloc(sizeof(void *) * (np > 0 ? np : 1));
    for (int k = 0; k < np; ++k) tabs[k] = make_tab(basis, g, p0 + k);
#pragma omp parallel for schedule(dynamic, 4)
    for (int iy = 0; iy < ny; ++iy) {
        for (int ix = 0; ix < nx; ++ix) {
            double acc = 0.0;
            for (int k = 0; k < np; ++k) {
                bwd_ctx b;
                b.row = sino + (size_t)(p0 + k) * nu;
                b.acc = 0.0;
                pixel_dispatch(basis, g, p0 + k, tabs[k], ix, iy, pad, emit_bwd, &b);
                acc += b.acc;
            }
            img[(size_t)iy * nx + ix] = acc;
        }
    }
    for (int k = 0; k < np; ++k) free(tabs[k]);
    free(tabs);
}

/* Footprint of one pixel in one view: returns the number of (bin, weight)
 * pairs (written up to cap). */
int or_footprint(int basis, const void *h, int p, int ix, int iy, int pad, int *u_out, double *w_out, int cap) {
    const ogeo *g = (const ogeo *)h;
    void *tab = make_tab(basis, g, p);
    list_ctx l = {u_out, w_out, 0, cap};
    pixel_dispatch(basis, g, p, tab, ix, iy, pad, emit_list, &l);
    free(tab);
    return l.n;
}

/* Number of footprint evaluations (sum of corner-range lengths) of views
 * [p0, p1): the reference's conv_count (layers/A_B1.cu:50) without skips. */
long long or_eval_count(int basis, const void *h, int p0, int p1) {
    const ogeo *g = (const ogeo *)h;
    fpt hw = (basis == OR_B1) ? (fpt)1.0f : (fpt)0.5f;
    long long n = 0;
#pragma omp parallel for reduction(+ : n)
    for (int p = p0; p < p1; ++p)
        for (int iy = 0; iy < g->ny; ++iy)
            for (int ix = 0; ix < g->nx; ++ix) {
                fpt us[4];
                int a, b;
                corner_range(g->pmi + 12 * p, ix, iy, hw, g->nu, us, &a, &b);
                n += b - a + 1;
            }
    return n;
}

/* ------------------------------------------------------------------------- */
/* FP64 exact strip integrals (independent of the footprint model).           */
/* value(pixel, bin) = average over the detector bin (uniform in detector     */
/* coordinate) of the line integral, along the ray source -> detector point, */
/* of the basis function: the pixel indicator (B0) or the bilinear hat (B1). */
/* ------------------------------------------------------------------------- */

/* chord length of the line through P with unit direction (ex, ey) inside the
 * box [x0,x1]x[y0,y1] */
static double chord_box(double Px, double Py, double ex, double ey, double x0, double x1, double y0,
                        double y1) {
    double tmin = -1e300, tmax = 1e300;
    if (fabs(ex) < 1e-300) {
        if (Px < x0 || Px > x1) return 0.0;
    } else {
        double ta = (x0 - Px) / ex, tb = (x1 - Px) / ex;
        if (ta > tb) { double t = ta; ta = tb; tb = t; }
        if (ta > tmin) tmin = ta;
        if (tb < tmax) tmax = tb;
    }
    if (fabs(ey) < 1e-300) {
        if (Py < y0 || Py > y1) return 0.0;
    } else {
        double ta = (y0 - Py) / ey, tb = (y1 - Py) / ey;
        if (ta > tb) { double t = ta; ta = tb; tb = t; }
        if (ta > tmin) tmin = ta;
        if (tb < tmax) tmax = tb;
    }
    return (tmax > tmin) ? (tmax - tmin) : 0.0;
}

/* line integral of hat(x)hat(y), hat(t) = max(0, 1 - |t|/h), centred at
 * (cx, cy), along the line P + s e (unit e): piecewise quadratic in s, split
 * at the crossings of x = cx + {-dx, 0, dx} and y = cy + {-dy, 0, dy} and
 * integrated exactly by Simpson's rule on every piece. */
static double hat_line(double Px, double Py, double ex, double ey, double cx, double cy, double dx,
                       double dy) {
    double br[8];
    int nb = 0;
    double lo = -1e300, hi = 1e300;
    /* support of the product: |x-cx| < dx and |y-cy| < dy */
    if (fabs(ex) < 1e-300) {
        if (fabs(Px - cx) >= dx) return 0.0;
    } else {
        double ta = (cx - dx - Px) / ex, tb = (cx + dx - Px) / ex;
        if (ta > tb) { double t = ta; ta = tb; tb = t; }
        if (ta > lo) lo = ta;
        if (tb < hi) hi = tb;
        br[nb++] = (cx - Px) / ex;
    }
    if (fabs(ey) < 1e-300) {
        if (fabs(Py - cy) >= dy) return 0.0;
    } else {
        double ta = (cy - dy - Py) / ey, tb = (cy + dy - Py) / ey;
        if (ta > tb) { double t = ta; ta = tb; tb = t; }
        if (ta > lo) lo = ta;
        if (tb < hi) hi = tb;
        br[nb++] = (cy - Py) / ey;
    }
    if (!(hi > lo)) return 0.0;
    double pts[4];
    int np_ = 0;
    pts[np_++] = lo;
    for (int i = 0; i < nb; ++i)
        if (br[i] > lo && br[i] < hi) pts[np_++] = br[i];
    pts[np_++] = hi;
    /* sort (at most 4 points) */
    for (int i = 1; i < np_; ++i)
        for (int j = i; j > 0 && pts[j] < pts[j - 1]; --j) { double t = pts[j]; pts[j] = pts[j - 1]; pts[j - 1] = t; }
    double sum = 0.0;
    for (int i = 0; i + 1 < np_; ++i) {
        double s0 = pts[i], s1 = pts[i + 1], sm = 0.5 * (s0 + s1);
        double f[3];
        double ss[3] = {s0, sm, s1};
        for (int k = 0; k < 3; ++k) {
            double x = Px + ss[k] * ex, y = Py + ss[k] * ey;
            double hx = 1.0 - fabs(x - cx) / dx, hy = 1.0 - fabs(y - cy) / dy;
            f[k] = (hx > 0 ? hx : 0) * (hy > 0 ? hy : 0);
        }
        sum += (s1 - s0) * (f[0] + 4 * f[1] + f[2]) / 6.0;
    }
    return sum;
}

static double ray_value(int basis, double px, double py, double qx, double qy, double cx, double cy,
                        double dx, double dy) {
    double ex = qx - px, ey = qy - py;
    double n = sqrt(ex * ex + ey * ey);
    ex /= n;
    ey /= n;
    if (basis == OR_B1) return hat_line(px, py, ex, ey, cx, cy, dx, dy);
    return chord_box(px, py, ex, ey, cx - 0.5 * dx, cx + 0.5 * dx, cy - 0.5 * dy, cy + 0.5 * dy);
}

/* detector parameter tau in [0,1] across bin u of the ray through (X, Y) */
static double tau_of(const ogeo *g, int p, int u, double X, double Y) {
    double uu = or_geo_u(g, p, X, Y);
    return uu - u;
}

/* 8-point Gauss-Legendre rule on [-1, 1] */
static const double GL8x[8] = {-0.96028985649753623, -0.79666647741362674, -0.52553240991632899,
                               -0.18343464249564980, 0.18343464249564980,  0.52553240991632899,
                               0.79666647741362674,  0.96028985649753623};
static const double GL8w[8] = {0.10122853629037626, 0.22238103445337447, 0.31370664587788729,
                               0.36268378337836198, 0.36268378337836198, 0.31370664587788729,
                               0.22238103445337447, 0.10122853629037626};

/* Exact average over bin u of view p of the basis line integral for pixel
 * (ix, iy), unit coefficient.  Gauss-Legendre (8 points) on the sub-intervals
 * of the bin between the rays through the basis function's kinks. */
double or_strip_exact(int basis, const void *h, int p, int u, int ix, int iy) {
    const ogeo *g = (const ogeo *)h;
    double dx = g->dx, dy = g->dy;
    double cx = (ix - 0.5 * (g->nx - 1)) * dx, cy = (iy - 0.5 * (g->ny - 1)) * dy;
    double px = g->src[2 * p], py = g->src[2 * p + 1];
    double pux = g->puv[2 * p], puy = g->puv[2 * p + 1];
    double ax = g->dtv[2 * p] + pux * (u - 0.5 * g->nu), ay = g->dtv[2 * p + 1] + puy * (u - 0.5 * g->nu);
    double cuts[16];
    int nc = 0;
    cuts[nc++] = 0.0;
    double hx = (basis == OR_B1) ? dx : 0.5 * dx, hy = (basis == OR_B1) ? dy : 0.5 * dy;
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j) {
            if (basis != OR_B1 && (i == 0 || j == 0)) continue;
            double t = tau_of(g, p, u, cx + i * hx, cy + j * hy);
            if (t > 0 && t < 1) cuts[nc++] = t;
        }
    cuts[nc++] = 1.0;
    for (int i = 1; i < nc; ++i)
        for (int j = i; j > 0 && cuts[j] < cuts[j - 1]; --j) { double t = cuts[j]; cuts[j] = cuts[j - 1]; cuts[j - 1] = t; }
    double total = 0.0;
    for (int k = 0; k + 1 < nc; ++k) {
        double t0 = cuts[k], t1 = cuts[k + 1];
        if (t1 <= t0) continue;
        double half = 0.5 * (t1 - t0), mid = 0.5 * (t1 + t0);
        for (int q = 0; q < 8; ++q) {
            double t = mid + half * GL8x[q];
            double qx = ax + pux * t, qy = ay + puy * t;
            total += half * GL8w[q] * ray_value(basis, px, py, qx, qy, cx, cy, dx, dy);
        }
    }
    return total;
}

/* Adaptive Simpson over the bin with Richardson correction (q + (q-v)/15),
 * the 1-D analogue of layers/Ref_projector.cu:156-214; returns the average
 * over the bin and writes the accumulated |q-v|/15 error estimate.  The
 * first four levels always subdivide: the integrand has kinks (rays through
 * the basis function's corners) that a coarse panel can miss entirely. */
static int ORACLE_SIMPSON_FORCED = 0; /* depth at or below which panels may stop */
static double simpson_rec(int basis, const ogeo *g, int p, double ax, double ay, double pux,
                          double puy, double cx, double cy, double t0, double t1, double f0,
                          double fm, double f1, double whole, double eps, int depth, double *err) {
    double dx = g->dx, dy = g->dy, px = g->src[2 * p], py = g->src[2 * p + 1];
    double tm = 0.5 * (t0 + t1);
    double tl = 0.5 * (t0 + tm), tr = 0.5 * (tm + t1);
    double fl = ray_value(basis, px, py, ax + pux * tl, ay + puy * tl, cx, cy, dx, dy);
    double fr = ray_value(basis, px, py, ax + pux * tr, ay + puy * tr, cx, cy, dx, dy);
    double left = (tm - t0) * (f0 + 4 * fl + fm) / 6.0;
    double right = (t1 - tm) * (fm + 4 * fr + f1) / 6.0;
    double q = left + right;
    if (depth <= 0 || (depth <= ORACLE_SIMPSON_FORCED && fabs(q - whole) <= 15.0 * eps)) {
        *err += fabs(q - whole) / 15.0;
        return q + (q - whole) / 15.0;
    }
    return simpson_rec(basis, g, p, ax, ay, pux, puy, cx, cy, t0, tm, f0, fl, fm, left, 0.5 * eps, depth - 1, err) +
           simpson_rec(basis, g, p, ax, ay, pux, puy, cx, cy, tm, t1, fm, fr, f1, right, 0.5 * eps, depth - 1, err);
}

#define ORACLE_SIMPSON_FORCED_LEVELS 4
double or_strip_simpson(int basis, const void *h, int p, int u, int ix, int iy, double eps, int depth,
                        double *err_out) {
    const ogeo *g = (const ogeo *)h;
    double dx = g->dx, dy = g->dy;
    double cx = (ix - 0.5 * (g->nx - 1)) * dx, cy = (iy - 0.5 * (g->ny - 1)) * dy;
    double px = g->src[2 * p], py = g->src[2 * p + 1];
    double pux = g->puv[2 * p], puy = g->puv[2 * p + 1];
    double ax = g->dtv[2 * p] + pux * (u - 0.5 * g->nu), ay = g->dtv[2 * p + 1] + puy * (u - 0.5 * g->nu);
    double f0 = ray_value(basis, px, py, ax, ay, cx, cy, dx, dy);
    double f1 = ray_value(basis, px, py, ax + pux, ay + puy, cx, cy, dx, dy);
    double fm = ray_value(basis, px, py, ax + 0.5 * pux, ay + 0.5 * puy, cx, cy, dx, dy);
    double whole = (f0 + 4 * fm + f1) / 6.0;
    double err = 0.0;
    ORACLE_SIMPSON_FORCED = depth - ORACLE_SIMPSON_FORCED_LEVELS;
    double v = simpson_rec(basis, g, p, ax, ay, pux, puy, cx, cy, 0.0, 1.0, f0, fm, f1, whole, eps, depth, &err);
    if (err_out) *err_out = err;
    return v;
}

/* Exact strip-integral A for a whole image over views [p0, p1), evaluated
 * over bins within the footprint's corner range widened by `pad`. */
void or_forward_exact(int basis, const void *h, const double *img, double *sino, int p0, int p1, int pad) {
    const ogeo *g = (const ogeo *)h;
    int nx = g->nx, ny = g->ny, nu = g->nu;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = p0; p < p1; ++p) {
        double *row = sino + (size_t)p * nu;
        memset(row, 0, sizeof(double) * nu);
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                double c = img[(size_t)iy * nx + ix];
                if (c == 0.0) continue;
                fpt us[4];
                int a, b;
                corner_range(g->pmi + 12 * p, ix, iy, (basis == OR_B1) ? (fpt)1.0f : (fpt)0.5f, nu, us, &a, &b);
                a = imax(0, a - pad);
                b = imin(nu - 1, b + pad);
                for (int u = a; u <= b; ++u) row[u] += c * or_strip_exact(basis, g, p, u, ix, iy);
            }
    }
}

/* ------------------------------------------------------------------------- */
/* Hexagonal lattices (layers/A_H0.cu, layers/A_Hl.cu).                       */
/* Lattice point (ic, il): il-th row, ic-th column, odd rows shifted by half  */
/* a column and one point shorter (A_H0.cu:172-183, A_Hl.cu:171-182).         */
/*   H0: hexagon indicator, circumradius dt, row pitch 1.5 dt, columns dt*sqrt3 */
/*   Hl: three-direction box spline (Courant hat), row pitch dt/2             */
/* Bin ranges use the six true hexagon vertices (the reference's H0 culling   */
/* projects vertices around the Hl row pitch, A_H0.cu:82-156 with            */
/* h0_init's dy = dt/2, and loses off-centre rows; fixed here).              */
/* H0 values are divided by 3: the reference sums three unit-integral rhombus */
/* splines (A_H0.cu:245), i.e. three times the hexagon's projection; its      */
/* commented matrix generator divides by 3 (A_H0.cu:438).                     */
/* ------------------------------------------------------------------------- */
#define OR_SQRT3 1.7320508075688772935274463415

/* FP64 truncated-power forward differences, spline_conv.hh:150-229 */
static double tp2(double y) { return y > 0 ? y : 0; }
static double tp3(double y) { return y > 0 ? y * y : 0; }
static double tp4(double y) { return y > 0 ? y * y * y : 0; }
static double fd2_2(double y, double h) { return (tp2(y) - tp2(y - h)) / h; }
static double fd2_1(double y, double h, double h2) { return (fd2_2(y, h) - fd2_2(y - h2, h)) / h2; }
static double fd3_3(double y, double h) { return (tp3(y) - tp3(y - h)) / h; }
static double fd3_2(double y, double h, double h2) { return (fd3_3(y, h) - fd3_3(y - h2, h)) / h2; }
static double fd3_1(double y, double h, double h2, double h3) { return (fd3_2(y, h, h2) - fd3_2(y - h3, h, h2)) / h3; }
static double fd4_4(double y, double h) { return (tp4(y) - tp4(y - h)) / h; }
static double fd4_3(double y, double h, double h2) { return (fd4_4(y, h) - fd4_4(y - h2, h)) / h2; }
static double fd4_2(double y, double h, double h2, double h3) { return (fd4_3(y, h, h2) - fd4_3(y - h3, h, h2)) / h3; }
static double fd4_1(double y, double h, double h2, double h3, double h4) { return (fd4_2(y, h, h2, h3) - fd4_2(y - h4, h, h2, h3)) / h4; }

/* Hl footprint: layers/A_Hl.cu:21-46 (directions with |h| <= 1e-8 dropped) */
static double hl_conv(double y, double h1, double h2, double h3, double h4) {
    double hs[4];
    int n = 0;
    if (fabs(h1) > 1e-8) hs[n++] = h1;
    if (fabs(h2) > 1e-8) hs[n++] = h2;
    if (fabs(h3) > 1e-8) hs[n++] = h3;
    if (fabs(h4) > 1e-8) hs[n++] = h4;
    switch (n) {
    case 2: return fd2_1(y, hs[0], hs[1]);
    case 3: return fd3_1(y, hs[0], hs[1], hs[2]) / 2.0;
    case 4: return fd4_1(y, hs[0], hs[1], hs[2], hs[3]) / 6.0;
    default: return 0.0; /* the reference asserts here */
    }
}

double or_hl_conv(double y, double h1, double h2, double h3, double h4) { return hl_conv(y, h1, h2, h3, h4); }

enum { OR_H0 = 3, OR_HL = 4 };

/* lattice point centre (physical) */
static void hex_centre(const ogeo *g, int basis, int ic, int il, double dt, fpt *ox, fpt *oy) {
    fpt fdt = (fpt)dt;
    fpt dx = fdt * (fpt)OR_SQRT3;
    fpt dy = (basis == OR_H0) ? fdt * (fpt)1.5f : fdt * (fpt)0.5f;
    fpt nl2 = (fpt)0.5f * (fpt)(g->ny - 1), nc2 = (fpt)0.5f * (fpt)(g->nx - 1), nc3 = (fpt)0.5f * (fpt)(g->nx - 2);
    *oy = ((fpt)il - nl2) * dy;
    *ox = ((il & 1) ? ((fpt)ic - nc3) : ((fpt)ic - nc2)) * dx;
}

/* bin range from the six hexagon vertices (circumradius dt, pointy-top) */
static void hex_range(const ogeo *g, int p, double ox, double oy, double dt, int *minu, int *maxu) {
    static const double vx[6] = {0.0, 0.5, 0.5, 0.0, -0.5, -0.5};
    static const double vy[6] = {-1.0, -0.5, 0.5, 1.0, 0.5, -0.5};
    double umin = 1e300, umax = -1e300;
    for (int k = 0; k < 6; ++k) {
        double u = or_geo_u(g, p, ox + vx[k] * OR_SQRT3 * dt, oy + vy[k] * dt);
        if (u < umin) umin = u;
        if (u > umax) umax = u;
    }
    *minu = imin(imax(0, (int)floor(umin)), g->nu - 1);
    *maxu = imin(imax(0, (int)floor(umax)), g->nu - 1);
}

/* H0 / Hl for one (lattice point, view): A_H0.cu:158-256, A_Hl.cu:168-266 */
static void hex_pixel(const ogeo *g, int basis, double dt, int p, int ic, int il, int pad, emit_fn emit,
                      void *ctx) {
    int nu = g->nu;
    fpt fdt = (fpt)dt;
    fpt dx = fdt * (fpt)OR_SQRT3;
    fpt dy = (basis == OR_H0) ? fdt * (fpt)1.5f : fdt * (fpt)0.5f;
    fpt px = (fpt)g->src[2 * p], py = (fpt)g->src[2 * p + 1];
    fpt pux = (fpt)g->puv[2 * p], puy = (fpt)g->puv[2 * p + 1];
    fpt dtx = (fpt)g->dtv[2 * p], dty = (fpt)g->dtv[2 * p + 1];
    fpt ox, oy;
    int minu, maxu;
    hex_centre(g, basis, ic, il, dt, &ox, &oy);
    hex_range(g, p, (double)ox, (double)oy, dt, &minu, &maxu);
    minu = imax(0, minu - pad);
    maxu = imin(nu - 1, maxu + pad);
    fpt sx = dtx + pux * ((fpt)minu - (fpt)0.5f * (fpt)nu + (fpt)0.5f);
    fpt sy = dty + puy * ((fpt)minu - (fpt)0.5f * (fpt)nu + (fpt)0.5f);
    for (int u = minu; u <= maxu; ++u, sx += pux, sy += puy) {
        fpt ax = sx - pux * (fpt)0.5f, ay = sy - puy * (fpt)0.5f;
        fpt bx = sx + pux * (fpt)0.5f, by = sy + puy * (fpt)0.5f;
        fpt num = ((ox - px) * (sx - px) + (oy - py) * (sy - py));
        fpt dena = ((ax - px) * (sx - px) + (ay - py) * (sy - py));
        fpt denb = ((bx - px) * (sx - px) + (by - py) * (sy - py));
        fpt ra = num / dena, rb = num / denb;
        fpt ax1 = px + (ax - px) * ra, ay1 = py + (ay - py) * ra;
        fpt bx1 = px + (bx - px) * rb, by1 = py + (by - py) * rb;
        fpt th = (fpt)atan2((double)(sx - px), (double)(sy - py));
        double va = dt * sin((double)th);
        double vb = dt * sin((double)th + 2.0 * OR_PI / 3.0);
        double vc = dt * sin((double)th + 4.0 * OR_PI / 3.0);
        double vblur = -sqrt((double)(bx1 - ax1) * (bx1 - ax1) + (double)(by1 - ay1) * (by1 - ay1));
        double y = sqrt((double)(ax1 - ox) * (ax1 - ox) + (double)(ay1 - oy) * (ay1 - oy));
        if ((ay - py) * (ox - px) < (oy - py) * (ax - px)) y = -y;
        double conv;
        if (basis == OR_H0) {
            conv = (double)b0_conv((fpt)y, (fpt)vblur, (fpt)va, (fpt)vb) + (double)b0_conv((fpt)y, (fpt)vblur, (fpt)vb, (fpt)vc) +
                   (double)b0_conv((fpt)y, (fpt)vblur, (fpt)vc, (fpt)va);
            conv /= 3.0;
        } else {
            conv = hl_conv(y, vblur, va, vb, vc);
        }
        if (conv == conv && conv > 0.0) emit(ctx, u, conv * (double)dx * (double)dy);
    }
}

/* Hex A / A^T.  Image layout img[il*nc + ic] (nc = g->nx columns, nl = g->ny rows);
 * the last column of odd rows is not a lattice point (A ignores it, A^T writes 0). */
void or_hex_forward(int basis, const void *h, double dt, const double *img, double *sino, int p0, int p1, int pad) {
    const ogeo *g = (const ogeo *)h;
    int nc = g->nx, nl = g->ny, nu = g->nu;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = p0; p < p1; ++p) {
        double *row = sino + (size_t)p * nu;
        memset(row, 0, sizeof(double) * nu);
        fwd_ctx f;
        f.row = row;
        for (int il = 0; il < nl; ++il)
            for (int ic = 0; ic < nc - (il & 1); ++ic) {
                f.c = img[(size_t)il * nc + ic];
                hex_pixel(g, basis, dt, p, ic, il, pad, emit_fwd, &f);
            }
    }
}

void or_hex_backward(int basis, const void *h, double dt, const double *sino, double *img, int p0, int p1, int pad) {
    const ogeo *g = (const ogeo *)h;
    int nc = g->nx, nl = g->ny, nu = g->nu;
#pragma omp parallel for schedule(dynamic, 4)
    for (int il = 0; il < nl; ++il)
        for (int ic = 0; ic < nc; ++ic) {
            double acc = 0.0;
            if (!((il & 1) && ic == nc - 1))
                for (int p = p0; p < p1; ++p) {
                    bwd_ctx b;
                    b.row = sino + (size_t)p * nu;
                    b.acc = 0.0;
                    hex_pixel(g, basis, dt, p, ic, il, pad, emit_bwd, &b);
                    acc += b.acc;
                }
            img[(size_t)il * nc + ic] = acc;
        }
}

int or_hex_footprint(int basis, const void *h, double dt, int p, int ic, int il, int pad, int *u_out, double *w_out, int cap) {
    const ogeo *g = (const ogeo *)h;
    list_ctx l = {u_out, w_out, 0, cap};
    hex_pixel(g, basis, dt, p, ic, il, pad, emit_list, &l);
    return l.n;
}

/* exact hexagon-basis line integrals (FP64) ------------------------------- */
/* pointy-top hexagon of circumradius dt: vertices at angles 90 + 60k deg;   */
/* edge normals at 60k deg, apothem dt*sqrt(3)/2.                            */
static double hex_norm(double x, double y, double dt) {
    double apo = 0.5 * OR_SQRT3 * dt, m = 0.0;
    for (int k = 0; k < 6; ++k) {
        double a = k * OR_PI / 3.0;
        double v = (x * cos(a) + y * sin(a)) / apo;
        if (v > m) m = v;
    }
    return m;
}

/* chord of the line P + s e through the hexagon centred at (cx, cy) */
static double hex_chord(double Px, double Py, double ex, double ey, double cx, double cy, double dt) {
    double apo = 0.5 * OR_SQRT3 * dt, lo = -1e300, hi = 1e300;
    for (int k = 0; k < 6; ++k) {
        double a = k * OR_PI / 3.0, nx = cos(a), ny = sin(a);
        double d0 = (Px - cx) * nx + (Py - cy) * ny, de = ex * nx + ey * ny;
        if (fabs(de) < 1e-300) {
            if (d0 > apo) return 0.0;
            continue;
        }
        double t = (apo - d0) / de;
        if (de > 0) { if (t < hi) hi = t; } else { if (t > lo) lo = t; }
    }
    return hi > lo ? hi - lo : 0.0;
}

/* line integral of the Courant hat max(0, 1 - |x|_hex): piecewise linear
 * along the line; split at crossings of the six spokes and integrate each
 * piece with the two-point trapezoid (exact for linear pieces). */
static double hex_hat_line(double Px, double Py, double ex, double ey, double cx, double cy, double dt) {
    double apo = 0.5 * OR_SQRT3 * dt, lo = -1e300, hi = 1e300;
    for (int k = 0; k < 6; ++k) {
        double a = k * OR_PI / 3.0, nx = cos(a), ny = sin(a);
        double d0 = (Px - cx) * nx + (Py - cy) * ny, de = ex * nx + ey * ny;
        if (fabs(de) < 1e-300) { if (d0 > apo) return 0.0; continue; }
        double t = (apo - d0) / de;
        if (de > 0) { if (t < hi) hi = t; } else { if (t > lo) lo = t; }
    }
    if (!(hi > lo)) return 0.0;
    double pts[16];
    int n = 0;
    pts[n++] = lo;
    for (int k = 0; k < 6; ++k) { /* spokes: rays from centre at 90 + 60k deg */
        double a = OR_PI / 2 + k * OR_PI / 3.0, sx = cos(a), sy = sin(a);
        double den = ex * sy - ey * sx;
        if (fabs(den) < 1e-300) continue;
        double t = ((cx - Px) * sy - (cy - Py) * sx) / den;
        if (t > lo && t < hi) pts[n++] = t;
    }
    pts[n++] = hi;
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && pts[j] < pts[j - 1]; --j) { double t = pts[j]; pts[j] = pts[j - 1]; pts[j - 1] = t; }
    double sum = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        double f0 = 1.0 - hex_norm(Px + pts[i] * ex - cx, Py + pts[i] * ey - cy, dt);
        double f1 = 1.0 - hex_norm(Px + pts[i + 1] * ex - cx, Py + pts[i + 1] * ey - cy, dt);
        sum += 0.5 * (pts[i + 1] - pts[i]) * ((f0 > 0 ? f0 : 0) + (f1 > 0 ? f1 : 0));
    }
    return sum;
}

/* exact average over bin u of the line integral of the lattice point's basis */
double or_hex_strip_exact(int basis, const void *h, double dt, int p, int u, int ic, int il) {
    const ogeo *g = (const ogeo *)h;
    double dy = (basis == OR_H0) ? 1.5 * dt : 0.5 * dt, dx = OR_SQRT3 * dt;
    double cy = (il - 0.5 * (g->ny - 1)) * dy;
    double cx = ((il & 1) ? (ic - 0.5 * (g->nx - 2)) : (ic - 0.5 * (g->nx - 1))) * dx;
    double px = g->src[2 * p], py = g->src[2 * p + 1];
    double pux = g->puv[2 * p], puy = g->puv[2 * p + 1];
    double ax = g->dtv[2 * p] + pux * (u - 0.5 * g->nu), ay = g->dtv[2 * p + 1] + puy * (u - 0.5 * g->nu);
    double cuts[16];
    int nc = 0;
    cuts[nc++] = 0.0;
    for (int k = -1; k < 6; ++k) {
        double X = cx, Y = cy;
        if (k >= 0) { double a = OR_PI / 2 + k * OR_PI / 3.0; X += dt * cos(a); Y += dt * sin(a); }
        double t = or_geo_u(g, p, X, Y) - u;
        if (t > 0 && t < 1) cuts[nc++] = t;
    }
    cuts[nc++] = 1.0;
    for (int i = 1; i < nc; ++i)
        for (int j = i; j > 0 && cuts[j] < cuts[j - 1]; --j) { double t = cuts[j]; cuts[j] = cuts[j - 1]; cuts[j - 1] = t; }
    double total = 0.0;
    for (int k = 0; k + 1 < nc; ++k) {
        double t0 = cuts[k], t1 = cuts[k + 1];
        if (t1 <= t0) continue;
        double half = 0.5 * (t1 - t0), mid = 0.5 * (t1 + t0);
        for (int q = 0; q < 8; ++q) {
            double t = mid + half * GL8x[q];
            double qx = ax + pux * t - px, qy = ay + puy * t - py;
            double nn = sqrt(qx * qx + qy * qy);
            double v = (basis == OR_H0) ? hex_chord(px, py, qx / nn, qy / nn, cx, cy, dt)
                                        : hex_hat_line(px, py, qx / nn, qy / nn, cx, cy, dt);
            total += half * GL8w[q] * v;
        }
    }
    return total;
}

/* ------------------------------------------------------------------------- */
/* 3-D cone beam, layers/A_B1_3D.cu:82-312: the in-plane B1 footprint of the  */
/* voxel's (ix, iy) column (the 2-D B1 formulas, A_B1_3D.cu:95-283) times an  */
/* axial trapezoid in detector rows, a path-length weight and the trapezoid's */
/* inverse area.  Volume vol[(iz*ny + iy)*nx + ix] (copied raw,               */
/* A_B1_3D.cu:716); projections proj[(p*nv + j)*nu + u] (A_B1_3D.cu:296,731). */
/* The source sits in the z = 0 plane (b1_3d_init passes no dz,               */
/* A_B1_3D.cu:702-703); the detector's v axis is +z with v = nv/2 on that     */
/* plane (utils/geo.cu:215, 267).                                             */
/* The reference computes the views at 3pi/4 and 7pi/4 through the singular   */
/* branch of utils/geo.cu:39-41 and A_B1_3D.cu:162-167, whose (x*1.5 - 1)     */
/* rescale holds only for nz = nv = 1; the closed-form map used here is the   */
/* projective map the other views use.  Its back-projector multiplies by the  */
/* zeroed output volume (A_B1_3D.cu:408, 436) and returns zeros; the          */
/* or_b1_3d_backward below is the transpose of the forward.                   */
/* ------------------------------------------------------------------------- */
typedef struct {
    ogeo *g;
    int nz, nv;
    double dz, dv;
    fpt *pmv; /* np*4: index-space v numerator row (ix, iy, iz, 1) */
} ogeo3;

void *or_geo3_create(int nx, int ny, int nz, int np, int nu, int nv, double dx, double dy, double dz,
                     double du, double dv, double lsd, double lso, const double *angles) {
    ogeo3 *t = (ogeo3 *)calloc(1, sizeof(ogeo3));
    t->g = (ogeo *)or_geo_create(nx, ny, np, nu, dx, dy, du, lsd, lso, angles, 0);
    t->nz = nz; t->nv = nv; t->dz = dz; t->dv = dv;
    t->pmv = (fpt *)malloc(sizeof(fpt) * 4 * np);
    const double fox = centre_offset(nx, dx), foy = centre_offset(ny, dy), foz = centre_offset(nz, dz);
    for (int p = 0; p < np; ++p) {
        const double *src = t->g->src + 2 * p, *dtv = t->g->dtv + 2 * p;
        double ex = dtv[0] - src[0], ey = dtv[1] - src[1];
        double L = sqrt(ex * ex + ey * ey);
        ex /= L;
        ey /= L;
        double d0 = -(src[0] * ex + src[1] * ey);
        double vc = nv / 2.0;
        /* v = vc + (L/dv) z / ((X - src).e), rows scaled so pm[11] == 1 */
        double r1[4] = {vc * ex / d0, vc * ey / d0, (L / dv) / d0, vc};
        fpt *q = t->pmv + 4 * p;
        q[0] = (fpt)(grid_step(dx) * r1[0]);
        q[1] = (fpt)(grid_step(dy) * r1[1]);
        q[2] = (fpt)(grid_step(dz) * r1[2]);
        q[3] = (fpt)(r1[3] - fox * r1[0] - foy * r1[1] - foz * r1[2]);
    }
    return t;
}

void or_geo3_destroy(void *h) {
    ogeo3 *t = (ogeo3 *)h;
    if (!t) return;
    or_geo_destroy(t->g);
    free(t->pmv);
    free(t);
}

/* In-plane B1 convs of column (ix, iy) over its clamped corner range, no
 * skip (A_B1_3D.cu:117-151, 261-277). */
static void b1_column_convs(const ogeo *g, int p, const fpt *vx, const fpt *vy, int ix, int iy,
                            int *minu_out, int *maxu_out, fpt *conv) {
    int nx = g->nx, ny = g->ny, nu = g->nu;
    fpt dx = (fpt)g->dx, dy = (fpt)g->dy;
    fpt px = (fpt)g->src[2 * p], py = (fpt)g->src[2 * p + 1];
    fpt pux = (fpt)g->puv[2 * p], puy = (fpt)g->puv[2 * p + 1];
    fpt dtx = (fpt)g->dtv[2 * p], dty = (fpt)g->dtv[2 * p + 1];
    fpt us[4];
    int minu, maxu;
    fpt ox = ((fpt)ix - (fpt)0.5f * (fpt)(nx - 1)) * dx;
    fpt oy = ((fpt)iy - (fpt)0.5f * (fpt)(ny - 1)) * dy;
    corner_range(g->pmi + 12 * p, ix, iy, (fpt)1.0f, nu, us, &minu, &maxu);
    fpt sx = dtx + pux * ((fpt)minu - (fpt)0.5f * (fpt)nu + (fpt)0.5f);
    fpt sy = dty + puy * ((fpt)minu - (fpt)0.5f * (fpt)nu + (fpt)0.5f);
    for (int u = minu; u <= maxu; ++u, sx += pux, sy += puy) {
        fpt ax = sx - pux * (fpt)0.5f, ay = sy - puy * (fpt)0.5f;
        fpt bx = sx + pux * (fpt)0.5f, by = sy + puy * (fpt)0.5f;
        fpt ra = ((ox - px) * (sx - px) + (oy - py) * (sy - py)) / ((ax - px) * (sx - px) + (ay - py) * (sy - py));
        fpt rb = ((ox - px) * (sx - px) + (oy - py) * (sy - py)) / ((bx - px) * (sx - px) + (by - py) * (sy - py));
        fpt ax1 = px + (ax - px) * ra, ay1 = py + (ay - py) * ra;
        fpt bx1 = px + (bx - px) * rb, by1 = py + (by - py) * rb;
        fpt x1 = (fpt)sqrt((double)((ax1 - ox) * (ax1 - ox) + (ay1 - oy) * (ay1 - oy)));
        fpt x2 = (fpt)sqrt((double)((bx1 - ox) * (bx1 - ox) + (by1 - oy) * (by1 - oy)));
        fpt sgn = ((ax1 - ox) * (bx1 - ox) + (ay1 - oy) * (by1 - oy) > 0) ? (fpt)1.0f : (fpt)-1.0f;
        conv[u - minu] = b1_conv(x1 * sgn, x2, fp_max(vx[u], vy[u]), fp_min(vx[u], vy[u]));
    }
    *minu_out = minu;
    *maxu_out = maxu;
}

/* Axial part of voxel (ix, iy, iz): sorted row coordinates of the trapezoid
 * (x index +-1 at the column's iy, z index +-1/2; A_B1_3D.cu:133-137,
 * 188-202), the clamped row range, and the factor
 * dx dy dz * weight * 2 / ((vs3 - vs0) + (vs2 - vs1)) (A_B1_3D.cu:209-210).
 * Returns 0 when the trapezoid lies past the last row (A_B1_3D.cu:204),
 * -1 for an empty row range, 1 otherwise. */
static int b1_3d_axial(const ogeo3 *t, int p, int ix, int iy, int iz, fpt *vs, int *minv, int *maxv,
                       fpt *cfac) {
    const ogeo *g = t->g;
    const fpt *pm = g->pmi + 12 * p, *pv = t->pmv + 4 * p;
    fpt sx1 = (fpt)ix - (fpt)1.0f, sx2 = (fpt)ix + (fpt)1.0f, fy = (fpt)iy;
    fpt v1 = pv[0] * sx1 + pv[1] * fy + pv[3];
    fpt v2 = pv[0] * sx2 + pv[1] * fy + pv[3];
    fpt v3 = pm[8] * sx1 + pm[9] * fy + pm[11];
    fpt v4 = pm[8] * sx2 + pm[9] * fy + pm[11];
    fpt z1 = (fpt)iz - (fpt)0.5f, z2 = (fpt)iz + (fpt)0.5f;
    vs[0] = (v1 + pv[2] * z1) / v3;
    vs[1] = (v1 + pv[2] * z2) / v3;
    vs[2] = (v2 + pv[2] * z1) / v4;
    vs[3] = (v2 + pv[2] * z2) / v4;
    /* sort4, layers/spline_conv.hh:24-39 (the network of corner_range) */
#define OR_SORT2(a, b) do { if ((a) > (b)) { fpt t_ = (b); (b) = (a); (a) = t_; } } while (0)
    OR_SORT2(vs[0], vs[1]);
    OR_SORT2(vs[2], vs[3]);
    OR_SORT2(vs[0], vs[2]);
    OR_SORT2(vs[1], vs[3]);
    OR_SORT2(vs[1], vs[2]);
#undef OR_SORT2
    int nv = t->nv;
    if (vs[0] >= (fpt)nv) return 0;
    *minv = imin(imax(0, fp_floor_int(vs[0])), nv - 1);
    *maxv = imin(imax(0, fp_floor_int(vs[3])), nv - 1);
    if (*minv > *maxv) return -1;
    fpt dx = (fpt)g->dx, dy = (fpt)g->dy, dz = (fpt)t->dz;
    fpt px = (fpt)g->src[2 * p], py = (fpt)g->src[2 * p + 1];
    fpt ox = ((fpt)ix - (fpt)0.5f * (fpt)(g->nx - 1)) * dx;
    fpt oy = ((fpt)iy - (fpt)0.5f * (fpt)(g->ny - 1)) * dy;
    fpt oz = ((fpt)iz - (fpt)0.5f * (fpt)(t->nz - 1)) * dz;
    fpt r2 = (ox - px) * (ox - px) + (oy - py) * (oy - py);
    fpt weight = (fpt)sqrt((double)((r2 + oz * oz) / r2));
    *cfac = (fpt)((double)dx * (double)dy * (double)dz * (double)weight *
                  (double)((fpt)2.0f / ((vs[3] - vs[0]) + (vs[2] - vs[1]))));
    return 1;
}

void or_b1_3d_forward(const void *h, const double *vol, double *proj, int p0, int p1) {
    const ogeo3 *t = (const ogeo3 *)h;
    const ogeo *g = t->g;
    int nx = g->nx, ny = g->ny, nz = t->nz, nu = g->nu, nv = t->nv;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = p0; p < p1; ++p) {
        double *pj = proj + (size_t)p * nv * nu;
        memset(pj, 0, sizeof(double) * (size_t)nv * nu);
        fpt *vx = (fpt *)malloc(sizeof(fpt) * nu), *vy = (fpt *)malloc(sizeof(fpt) * nu);
        fpt *conv = (fpt *)malloc(sizeof(fpt) * nu);
        b1_vectors(g, p, vx, vy);
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                int minu, maxu, minv, maxv;
                int any = 0;
                for (int iz = 0; iz < nz && !any; ++iz) any = vol[((size_t)iz * ny + iy) * nx + ix] != 0.0;
                if (!any) continue;
                b1_column_convs(g, p, vx, vy, ix, iy, &minu, &maxu, conv);
                for (int iz = 0; iz < nz; ++iz) {
                    double x = vol[((size_t)iz * ny + iy) * nx + ix];
                    fpt vs[4], cf;
                    int r = b1_3d_axial(t, p, ix, iy, iz, vs, &minv, &maxv, &cf);
                    if (r == 0) break;
                    if (r < 0 || x == 0.0) continue;
                    fpt c = (fpt)(x * (double)cf);
                    for (int u = minu; u <= maxu; ++u) {
                        double val = (double)(conv[u - minu] * c);
                        if (val != val) continue;
                        for (int j = minv; j <= maxv; ++j) {
                            fpt gm = sf_gamma((fpt)j, (fpt)j + (fpt)1.0f, vs);
                            if (gm > EPS7) pj[(size_t)j * nu + u] += val * (double)gm;
                        }
                    }
                }
            }
        free(vx); free(vy); free(conv);
    }
}

/* Transpose of or_b1_3d_forward (the reference's own back-projector returns
 * zeros, see above): vol[voxel] = sum over views, bins and rows of
 * conv * cfac * gamma * proj. */
void or_b1_3d_backward(const void *h, const double *proj, double *vol, int p0, int p1) {
    const ogeo3 *t = (const ogeo3 *)h;
    const ogeo *g = t->g;
    int nx = g->nx, ny = g->ny, nz = t->nz, nu = g->nu, nv = t->nv, np = g->np;
    fpt *vxs = (fpt *)malloc(sizeof(fpt) * (size_t)np * nu), *vys = (fpt *)malloc(sizeof(fpt) * (size_t)np * nu);
    for (int p = p0; p < p1; ++p) b1_vectors(g, p, vxs + (size_t)p * nu, vys + (size_t)p * nu);
#pragma omp parallel for schedule(dynamic, 1)
    for (int col = 0; col < nx * ny; ++col) {
        int ix = col % nx, iy = col / nx;
        fpt *conv = (fpt *)malloc(sizeof(fpt) * nu);
        for (int iz = 0; iz < nz; ++iz) vol[((size_t)iz * ny + iy) * nx + ix] = 0.0;
        for (int p = p0; p < p1; ++p) {
            int minu, maxu, minv, maxv;
            const double *pj = proj + (size_t)p * nv * nu;
            b1_column_convs(g, p, vxs + (size_t)p * nu, vys + (size_t)p * nu, ix, iy, &minu, &maxu, conv);
            for (int iz = 0; iz < nz; ++iz) {
                fpt vs[4], cf;
                int r = b1_3d_axial(t, p, ix, iy, iz, vs, &minv, &maxv, &cf);
                if (r == 0) break;
                if (r < 0) continue;
                double acc = 0.0;
                for (int u = minu; u <= maxu; ++u) {
                    double cu = (double)conv[u - minu] * (double)cf;
                    if (cu != cu) continue;
                    for (int j = minv; j <= maxv; ++j) {
                        fpt gm = sf_gamma((fpt)j, (fpt)j + (fpt)1.0f, vs);
                        if (gm > EPS7) acc += cu * (double)gm * pj[(size_t)j * nu + u];
                    }
                }
                vol[((size_t)iz * ny + iy) * nx + ix] += acc;
            }
        }
        free(conv);
    }
    free(vxs);
    free(vys);
}

/* Line integral of hat(x)hat(y)box(z) (hat half-widths dx, dy; box height
 * dz) centred at c along P + s e (unit e): the support is an interval, the
 * integrand piecewise quadratic with kinks where x = cx or y = cy; Simpson
 * on every piece is exact. */
static double hat_box_line3(const double *P, const double *e, const double *c, double dx, double dy,
                            double dz) {
    double lo = -1e300, hi = 1e300, br[2];
    int nb = 0;
    const double half[3] = {dx, dy, 0.5 * dz};
    for (int k = 0; k < 3; ++k) {
        if (fabs(e[k]) < 1e-300) {
            if (fabs(P[k] - c[k]) >= half[k]) return 0.0;
            continue;
        }
        double ta = (c[k] - half[k] - P[k]) / e[k], tb = (c[k] + half[k] - P[k]) / e[k];
        if (ta > tb) { double tt = ta; ta = tb; tb = tt; }
        if (ta > lo) lo = ta;
        if (tb < hi) hi = tb;
        if (k < 2) br[nb++] = (c[k] - P[k]) / e[k];
    }
    if (!(hi > lo)) return 0.0;
    double pts[4];
    int n = 0;
    pts[n++] = lo;
    for (int i = 0; i < nb; ++i)
        if (br[i] > lo && br[i] < hi) pts[n++] = br[i];
    pts[n++] = hi;
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && pts[j] < pts[j - 1]; --j) { double tt = pts[j]; pts[j] = pts[j - 1]; pts[j - 1] = tt; }
    double sum = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        double s0 = pts[i], s1 = pts[i + 1], ss[3] = {s0, 0.5 * (s0 + s1), s1}, f[3];
        for (int k = 0; k < 3; ++k) {
            double hx = 1.0 - fabs(P[0] + ss[k] * e[0] - c[0]) / dx;
            double hy = 1.0 - fabs(P[1] + ss[k] * e[1] - c[1]) / dy;
            f[k] = (hx > 0 ? hx : 0) * (hy > 0 ? hy : 0);
        }
        sum += (s1 - s0) * (f[0] + 4 * f[1] + f[2]) / 6.0;
    }
    return sum;
}

/* FP64 mean over detector cell (u, j) of the line integral of voxel
 * (ix, iy, iz)'s basis along source -> detector-point rays: composite
 * Gauss-Legendre, `panels` x `panels` panels of 8 x 8 points. */
double or_b1_3d_cell_exact(const void *h, int p, int u, int j, int ix, int iy, int iz, int panels) {
    const ogeo3 *t = (const ogeo3 *)h;
    const ogeo *g = t->g;
    double c[3] = {(ix - 0.5 * (g->nx - 1)) * g->dx, (iy - 0.5 * (g->ny - 1)) * g->dy,
                   (iz - 0.5 * (t->nz - 1)) * t->dz};
    double P[3] = {g->src[2 * p], g->src[2 * p + 1], 0.0};
    double pux = g->puv[2 * p], puy = g->puv[2 * p + 1];
    double total = 0.0;
    for (int a = 0; a < panels; ++a)
        for (int qa = 0; qa < 8; ++qa) {
            double s = (a + 0.5 + 0.5 * GL8x[qa]) / panels, ws = GL8w[qa] * 0.5 / panels;
            double qx = g->dtv[2 * p] + pux * (u + s - 0.5 * g->nu);
            double qy = g->dtv[2 * p + 1] + puy * (u + s - 0.5 * g->nu);
            for (int b = 0; b < panels; ++b)
                for (int qb = 0; qb < 8; ++qb) {
                    double tt = (b + 0.5 + 0.5 * GL8x[qb]) / panels, wt = GL8w[qb] * 0.5 / panels;
                    double qz = t->dv * (j + tt - 0.5 * t->nv);
                    double e[3] = {qx - P[0], qy - P[1], qz - P[2]};
                    double nn = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                    e[0] /= nn; e[1] /= nn; e[2] /= nn;
                    total += ws * wt * hat_box_line3(P, e, c, g->dx, g->dy, t->dz);
                }
        }
    return total;
}

/* ------------------------------------------------------------------------- */
/* Explicit system matrices: pixel_matrix_generation (B0, layers/A_B0.cu:36-  */
/* 121 kernel, :449-492 host) and bilinear_matrix_generation (B1,             */
/* A_B1.cu:355-427, :545-593).  Views from geo_init_example(lsd, lso, 0,      */
/* 2 pi (np-1)/np) (utils/geo.cu:209-230: equal steps, no 45-degree           */
/* snapping); every (bin, pixel) pair passing the kernels' two culls with a   */
/* positive footprint becomes a COO triplet (row p*nu + u, col i*ny + j, conv)*/
/* in the order view, bin, i, j.  Returns the triplet count, or -1 when more  */
/* than buffer_size triplets would be written (the reference asserts).        */
/* B1, fixed = 0: the reference's kernel, which averages over [y, y + vblur]  */
/* (A_B1.cu:402-411) and so returns bin u-1's footprint in row u (its own     */
/* projector, A_B1.cu:182-197, and its quadrature generator                   */
/* python/matrix_gen/A_B1_ref_project.py agree with each other, not with it). */
/* fixed = 1: the projector's interval (x1 sgn, x2) -- the matrix of          */
/* b1_forward_projection.  B0 ignores `fixed`.                                */
/* ------------------------------------------------------------------------- */
int or_matrix_generation(int basis, int np, int nu, int nx, int ny, double dx, double dy, double du,
                         double lsd, double lso, double *spmat, long long buffer_size, int fixed) {
    const double PI_ = 3.14159265358979323846;
    double end = PI_ * 2 * (np - 1) / np;
    double dangle = np > 1 ? end / (np - 1) : 0.0;
    double nx2 = 0.5 * (nx - 1), ny2 = 0.5 * (ny - 1);
    long long idx = 0;
    double *row = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    for (int p = 0; p < np; ++p) {
        double beta = p * dangle, c = cos(beta), sn = sin(beta);
        double px = -lso * c, py = -lso * sn;
        double dtx = (lsd - lso) * c, dty = (lsd - lso) * sn;
        double pux = -du * sn, puy = du * c;
        for (int u = 0; u < nu; ++u) {
            double sx = dtx + pux * (u - 0.5 * nu + 0.5), sy = dty + puy * (u - 0.5 * nu + 0.5);
            double ax = sx - pux * 0.5, ay = sy - puy * 0.5, bx = sx + pux * 0.5, by = sy + puy * 0.5;
            double theta = atan2(fabs(px - sx), fabs(py - sy));
            double sint = sin(theta), cost = cos(theta);
            double vx, vy;
            if (basis == OR_B0) {
                vx = fabs(dx * cost);
                vy = fabs(dy * sint);
                if (py < sy) vx = -vx;
                if (px < sx) vy = -vy;
            } else {
                vx = dx * cost;
                vy = dy * sint;
            }
            for (int i = 0; i < nx; ++i)
                for (int j = 0; j < ny; ++j) {
                    double ox = (i - nx2) * dx, oy = (j - ny2) * dy;
                    double num = (ox - px) * (sx - px) + (oy - py) * (sy - py);
                    double ra = num / ((ax - px) * (sx - px) + (ay - py) * (sy - py));
                    double rb = num / ((bx - px) * (sx - px) + (by - py) * (sy - py));
                    double ax1 = px + (ax - px) * ra, ay1 = py + (ay - py) * ra;
                    double bx1 = px + (bx - px) * rb, by1 = py + (by - py) * rb;
                    row[i * ny + j] = 0.0;
                    if (fabs(ax1 - ox) > dx && fabs(bx1 - ox) > dx && fabs(ay1 - oy) > dy && fabs(by1 - oy) > dy)
                        continue;
                    fpt conv;
                    if (basis == OR_B0) {
                        double om = -0.5 * dx * (px - sx) + 0.5 * dy * (py - sy);
                        double l2 = (sx - px) * (sx - px) + (sy - py) * (sy - py);
                        double zx = ox - 0.5 * dx - om * (px - sx) / l2, zy = oy + 0.5 * dy - om * (py - sy) / l2;
                        double vblur = -sqrt((bx1 - ax1) * (bx1 - ax1) + (by1 - ay1) * (by1 - ay1));
                        double y = sqrt((ax1 - zx) * (ax1 - zx) + (ay1 - zy) * (ay1 - zy));
                        if (y > fabs(vx) + fabs(vy) + fabs(vblur)) continue;
                        if ((ay - py) * (zx - px) < (zy - py) * (ax - px)) y = -y;
                        conv = b0_conv((fpt)y, (fpt)vx, (fpt)vy, (fpt)vblur);
                    } else {
                        double vblur = sqrt((bx1 - ax1) * (bx1 - ax1) + (by1 - ay1) * (by1 - ay1));
                        double y = sqrt((ax1 - ox) * (ax1 - ox) + (ay1 - oy) * (ay1 - oy));
                        if (y > 2 * fabs(vx) + 2 * fabs(vy) + fabs(vblur)) continue;
                        if ((ay - py) * (ox - px) > (oy - py) * (ax - px)) y = -y;
                        fpt a = (fpt)(vx > vy ? vx : vy), b = (fpt)(vx < vy ? vx : vy);
                        if (fixed) {
                            double x2 = sqrt((bx1 - ox) * (bx1 - ox) + (by1 - oy) * (by1 - oy));
                            double sg = ((ax1 - ox) * (bx1 - ox) + (ay1 - oy) * (by1 - oy) > 0) ? 1.0 : -1.0;
                            double x1 = sqrt((ax1 - ox) * (ax1 - ox) + (ay1 - oy) * (ay1 - oy));
                            conv = b1_conv((fpt)(x1 * sg), (fpt)x2, a, b);
                        } else {
                            conv = b1_conv((fpt)y, (fpt)(y + vblur), a, b);
                        }
                    }
                    if (conv > 0) row[i * ny + j] = (double)conv;
                }
            for (int k = 0; k < nx * ny; ++k)
                if (row[k] != 0.0) {
                    if (idx == buffer_size) {
                        free(row);
                        return -1;
                    }
                    spmat[3 * idx] = (double)p * nu + u;
                    spmat[3 * idx + 1] = (double)k;
                    spmat[3 * idx + 2] = row[k];
                    ++idx;
                }
        }
    }
    free(row);
    return (int)idx;
}

/* ------------------------------------------------------------------------- */
/* 3-D separable footprints, layers/SF_projector.cu:81-234: the voxel's u     */
/* trapezoid (corners ix +- 1/2, iy +- 1/2) times a rectangle (SF-TR: the     */
/* centre column at iz +- 1/2) or a trapezoid (SF-TT: x index +- 1/2 at iy,   */
/* iz +- 1/2) in v, on detector cells centred on integer coordinates          */
/* ([i - 1/2, i + 1/2], SF_projector.cu:209-216), amplitude                   */
/* lsd^2 dx dy dz / (du dv) / r^2 times the inverse areas of the two          */
/* footprints (SF_projector.cu:193-198, 364).  Fixed against the reference:   */
/* the voxel volume dx dy dz (:364 has dy dz, the volume only when dx = 1;    */
/* against exact chord integrals its mass is off by 1/dx); the value          */
/* carries the voxel and the amplitude (its forward adds only the footprint   */
/* product f2, :230-231, leaving C unused), SF-TT's v area is                 */
/* 2 / ((v3 - v0) + (v2 - v1)) (:195 reads 2 / (v3 - v0) + (v2 - v1)), r is   */
/* measured from the voxel centre ((nx - 1) / 2 offsets; :103 uses nx / 2 in  */
/* integers), and cells outside the detector are dropped (:228 wraps u past   */
/* a row end into the next row).  Same volume / projection layouts as         */
/* or_b1_3d_*; the transpose is or_sf3d_backward.                             */
/* ------------------------------------------------------------------------- */
static int sf3d_voxel(const ogeo3 *t, int p, int ix, int iy, int iz, int trap_v, fpt *us, fpt *vs, int *u0, int *u1,
                      int *v0, int *v1, fpt *amp) {
    const ogeo *g = t->g;
    const fpt *pm = g->pmi + 12 * p, *pv = t->pmv + 4 * p;
    int nu = g->nu, nv = t->nv, dummy0, dummy1;
    corner_range(pm, ix, iy, (fpt)0.5f, nu, us, &dummy0, &dummy1); /* sorted corner projections */
    fpt z1 = (fpt)iz - (fpt)0.5f, z2 = (fpt)iz + (fpt)0.5f;
    fpt av;
    if (!trap_v) {
        fpt num = pv[0] * (fpt)ix + pv[1] * (fpt)iy + pv[3];
        fpt den = pm[8] * (fpt)ix + pm[9] * (fpt)iy + pm[11];
        vs[0] = (num + pv[2] * z1) / den;
        vs[1] = (num + pv[2] * z2) / den;
        if (ceil((double)vs[0] - 0.5) >= nv) return 0;
        av = (fpt)1.0f / (vs[1] - vs[0]);
        *v0 = (int)ceil((double)vs[0] - 0.5);
        *v1 = (int)ceil((double)vs[1] - 0.5);
    } else {
        fpt sx1 = (fpt)ix - (fpt)0.5f, sx2 = (fpt)ix + (fpt)0.5f, fy = (fpt)iy;
        fpt n1 = pv[0] * sx1 + pv[1] * fy + pv[3], n2 = pv[0] * sx2 + pv[1] * fy + pv[3];
        fpt d1 = pm[8] * sx1 + pm[9] * fy + pm[11], d2 = pm[8] * sx2 + pm[9] * fy + pm[11];
        vs[0] = (n1 + pv[2] * z1) / d1;
        vs[1] = (n1 + pv[2] * z2) / d1;
        vs[2] = (n2 + pv[2] * z1) / d2;
        vs[3] = (n2 + pv[2] * z2) / d2;
#define OR_SORT2(a, b) do { if ((a) > (b)) { fpt t_ = (b); (b) = (a); (a) = t_; } } while (0)
        OR_SORT2(vs[0], vs[1]);
        OR_SORT2(vs[2], vs[3]);
        OR_SORT2(vs[0], vs[2]);
        OR_SORT2(vs[1], vs[3]);
        OR_SORT2(vs[1], vs[2]);
#undef OR_SORT2
        av = (fpt)2.0f / ((vs[3] - vs[0]) + (vs[2] - vs[1]));
        *v0 = (int)ceil((double)vs[0] - 0.5);
        *v1 = (int)ceil((double)vs[3] - 0.5);
    }
    *u0 = (int)ceil((double)us[0] - 0.5);
    *u1 = (int)ceil((double)us[3] - 0.5);
    if (*v1 < 0 || *v0 >= nv || *u1 < 0 || *u0 >= nu) return -1;
    *u0 = imax(*u0, 0); *u1 = imin(*u1, nu - 1);
    *v0 = imax(*v0, 0); *v1 = imin(*v1, nv - 1);
    double dx = g->dx, dy = g->dy, dz = t->dz;
    double ox = (ix - 0.5 * (g->nx - 1)) * dx - g->src[2 * p], oy = (iy - 0.5 * (g->ny - 1)) * dy - g->src[2 * p + 1];
    double oz = (iz - 0.5 * (t->nz - 1)) * dz;
    double lsd = g->lsd;
    double fac = lsd * lsd * dx * dy * dz / (g->du * t->dv);
    double w = 1.0 / (ox * ox + oy * oy + oz * oz);
    *amp = (fpt)(w * fac * (double)((fpt)2.0f / ((us[3] - us[0]) + (us[2] - us[1]))) * (double)av);
    return 1;
}

static fpt sf3d_row_weight(int trap_v, const fpt *vs, int j) {
    fpt s1 = (fpt)j - (fpt)0.5f, s2 = (fpt)j + (fpt)0.5f;
    if (trap_v) return sf_gamma(s1, s2, vs);
    fpt a = s2 < vs[1] ? s2 : vs[1], b = s1 > vs[0] ? s1 : vs[0];
    return a - b > 0 ? a - b : (fpt)0.0f;
}

void or_sf3d_forward(const void *h, int trap_v, const double *vol, double *proj, int p0, int p1) {
    const ogeo3 *t = (const ogeo3 *)h;
    const ogeo *g = t->g;
    int nx = g->nx, ny = g->ny, nz = t->nz, nu = g->nu, nv = t->nv;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = p0; p < p1; ++p) {
        double *pj = proj + (size_t)p * nv * nu;
        memset(pj, 0, sizeof(double) * (size_t)nv * nu);
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    double x = vol[((size_t)iz * ny + iy) * nx + ix];
                    if (x == 0.0) continue;
                    fpt us[4], vs[4], amp;
                    int u0, u1, v0, v1;
                    if (sf3d_voxel(t, p, ix, iy, iz, trap_v, us, vs, &u0, &u1, &v0, &v1, &amp) <= 0) continue;
                    for (int u = u0; u <= u1; ++u) {
                        fpt gu = sf_gamma((fpt)u - (fpt)0.5f, (fpt)u + (fpt)0.5f, us);
                        for (int j = v0; j <= v1; ++j) {
                            double f = (double)amp * (double)gu * (double)sf3d_row_weight(trap_v, vs, j);
                            if (f == f) pj[(size_t)j * nu + u] += x * f;
                        }
                    }
                }
    }
}

void or_sf3d_backward(const void *h, int trap_v, const double *proj, double *vol, int p0, int p1) {
    const ogeo3 *t = (const ogeo3 *)h;
    const ogeo *g = t->g;
    int nx = g->nx, ny = g->ny, nz = t->nz, nu = g->nu, nv = t->nv;
#pragma omp parallel for schedule(dynamic, 1)
    for (int col = 0; col < nx * ny; ++col) {
        int ix = col % nx, iy = col / nx;
        for (int iz = 0; iz < nz; ++iz) {
            double acc = 0.0;
            for (int p = p0; p < p1; ++p) {
                const double *pj = proj + (size_t)p * nv * nu;
                fpt us[4], vs[4], amp;
                int u0, u1, v0, v1;
                if (sf3d_voxel(t, p, ix, iy, iz, trap_v, us, vs, &u0, &u1, &v0, &v1, &amp) <= 0) continue;
                for (int u = u0; u <= u1; ++u) {
                    fpt gu = sf_gamma((fpt)u - (fpt)0.5f, (fpt)u + (fpt)0.5f, us);
                    for (int j = v0; j <= v1; ++j) {
                        double f = (double)amp * (double)gu * (double)sf3d_row_weight(trap_v, vs, j);
                        if (f == f) acc += f * pj[(size_t)j * nu + u];
                    }
                }
            }
            vol[((size_t)iz * ny + iy) * nx + ix] = acc;
        }
    }
}

/* chord length of the line P + s e (unit e) through the box [lo, hi] */
static double box_chord3(const double *P, const double *e, const double *lo, const double *hi) {
    double a = -1e300, b = 1e300;
    for (int k = 0; k < 3; ++k) {
        if (fabs(e[k]) < 1e-300) {
            if (P[k] < lo[k] || P[k] > hi[k]) return 0.0;
            continue;
        }
        double t0 = (lo[k] - P[k]) / e[k], t1 = (hi[k] - P[k]) / e[k];
        if (t0 > t1) { double tt = t0; t0 = t1; t1 = tt; }
        if (t0 > a) a = t0;
        if (t1 < b) b = t1;
    }
    return b > a ? b - a : 0.0;
}

/* FP64 mean over SF detector cell (u, j) = [u - 1/2, u + 1/2] x [j - 1/2,
 * j + 1/2] of the chord length of source -> detector-point rays through voxel
 * (ix, iy, iz) (composite Gauss-Legendre, panels x panels of 8 x 8). */
double or_sf3d_cell_exact(const void *h, int p, int u, int j, int ix, int iy, int iz, int panels) {
    const ogeo3 *t = (const ogeo3 *)h;
    const ogeo *g = t->g;
    double c[3] = {(ix - 0.5 * (g->nx - 1)) * g->dx, (iy - 0.5 * (g->ny - 1)) * g->dy,
                   (iz - 0.5 * (t->nz - 1)) * t->dz};
    double lo[3] = {c[0] - 0.5 * g->dx, c[1] - 0.5 * g->dy, c[2] - 0.5 * t->dz};
    double hi[3] = {c[0] + 0.5 * g->dx, c[1] + 0.5 * g->dy, c[2] + 0.5 * t->dz};
    double P[3] = {g->src[2 * p], g->src[2 * p + 1], 0.0};
    double pux = g->puv[2 * p], puy = g->puv[2 * p + 1];
    double total = 0.0;
    for (int a = 0; a < panels; ++a)
        for (int qa = 0; qa < 8; ++qa) {
            double s = (a + 0.5 + 0.5 * GL8x[qa]) / panels - 0.5, ws = GL8w[qa] * 0.5 / panels;
            double qx = g->dtv[2 * p] + pux * (u + s - 0.5 * g->nu);
            double qy = g->dtv[2 * p + 1] + puy * (u + s - 0.5 * g->nu);
            for (int b = 0; b < panels; ++b)
                for (int qb = 0; qb < 8; ++qb) {
                    double tt = (b + 0.5 + 0.5 * GL8x[qb]) / panels - 0.5, wt = GL8w[qb] * 0.5 / panels;
                    double qz = t->dv * (j + tt - 0.5 * t->nv);
                    double e[3] = {qx - P[0], qy - P[1], qz - P[2]};
                    double nn = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                    e[0] /= nn; e[1] /= nn; e[2] /= nn;
                    total += ws * wt * box_chord3(P, e, lo, hi);
                }
        }
    return total;
}

int or_precision_bits(void) { return (int)(sizeof(fpt) * 8); }
