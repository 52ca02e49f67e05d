/*
 * sgdct.h -- C ABI of the MI355X (gfx950) box-spline footprint projectors.
 *
 * Drop-in for the extern "C" surface of hqpwca/SGD's projector layers.  The
 * reference-named entry points keep the reference's argument order, meaning
 * and host layouts:
 *   image     x[ix*ny + iy]  (FP64, host)       layers/A_B1.cu:621-623
 *   sinogram  b[p*nu + u]    (FP64, host)       layers/A_B1.cu:639-641
 *   A overwrites b, A^T overwrites x.
 * Differences from the reference, all additive:
 *   - dz / drho are explicit nullable pointers (the reference relies on C++
 *     default arguments that do not exist across a C ABI, A_B1.hh:24).
 *   - ny is honoured (the reference reads ny = nx, layers/A_B1.cu:607).
 *   - errors are negative return codes, never exceptions; sgdct_last_error()
 *     describes the last one on the calling thread.
 *   - a destroy function, and FP32 device-pointer entry points on a caller
 *     stream for callers that keep data on the GPU.
 */
#ifndef SGDCT_H
#define SGDCT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sgdct_projector sgdct_projector;

/* Error codes (negative). */
#define SGDCT_OK 0
#define SGDCT_EINVAL (-1)   /* bad argument or shape */
#define SGDCT_EDEVICE (-2)  /* HIP runtime error (no device, launch failure) */
#define SGDCT_ENOMEM (-3)   /* device allocation failed */
#define SGDCT_EUNSUP (-4)   /* basis / operation not supported */

/* ---- reference-named entry points -------------------------------------- */

/* layers/A_B0.hh:24 (b0_init), A_B0.cu:504 (forward), A_B0.cu:543 (backward) */
sgdct_projector *b0_init(int nx, int ny, int np, int nu, double dx, double dy, double du,
                         double lsd, double lso, const double *angles, const double *dz,
                         const double *drho);
int b0_forward_projection(double *b, double *x, sgdct_projector *h);
int b0_backward_projection(double *b, double *x, sgdct_projector *h);

/* layers/A_B1.hh:24 (b1_init), A_B1.cu:605 (forward), A_B1.cu:646 (backward) */
sgdct_projector *b1_init(int nx, int ny, int np, int nu, double dx, double dy, double du,
                         double lsd, double lso, const double *angles, const double *dz,
                         const double *drho);
int b1_forward_projection(double *b, double *x, sgdct_projector *h);
int b1_backward_projection(double *b, double *x, sgdct_projector *h);

/* layers/A_SF.hh (SF_init), A_SF.cu:396 (init), :406 (forward), :447 (backward) */
sgdct_projector *SF_init(int nx, int ny, int np, int nu, double dx, double dy, double du,
                         double lsd, double lso, const double *angles);
int SF_forward_projection(double *b, double *x, sgdct_projector *h);
int SF_backward_projection(double *b, double *x, sgdct_projector *h);

/* layers/A_H0.hh:24 / A_H0.cu:509 (h0_init), :556 (forward), :601 (backward);
 * nl rows x nc columns of a hexagonal lattice, host image x[ic*nl + il]; the
 * last column of odd rows is not a lattice point (ignored by A, zeroed by A^T).
 * Fixed against the reference: bin ranges from the true hexagon vertices
 * (A_H0.cu:82-156 uses the Hl row pitch), values divided by 3 (A_H0.cu:245
 * sums three unit-integral rhombi; its matrix generator divides, :438). */
sgdct_projector *h0_init(int nl, int nc, int np, int nu, double dt, double du, double lsd,
                         double lso, const double *angles, const double *dz, const double *drho);
int h0_forward_projection(double *b, double *x, sgdct_projector *h);
int h0_backward_projection(double *b, double *x, sgdct_projector *h);

/* layers/A_Hl.hh:24 / A_Hl.cu:521 (hl_init), :569 (forward), :614 (backward);
 * three-direction box-spline basis on the row pitch dt/2 lattice. */
sgdct_projector *hl_init(int nl, int nc, int np, int nu, double dt, double du, double lsd,
                         double lso, const double *angles, const double *dz, const double *drho);
int hl_forward_projection(double *b, double *x, sgdct_projector *h);
int hl_backward_projection(double *b, double *x, sgdct_projector *h);

/* layers/A_B0.cu:449-492 (pixel_matrix_generation) and A_B1.cu:545-593
 * (bilinear_matrix_generation): the explicit system matrix of B0 / B1 for np
 * equally spaced views over [0, 2 pi (np-1)/np], as COO triplets
 * spmat[3k..3k+2] = (p*nu + u, i*ny + j, conv) in the order view, bin, i, j,
 * kept where conv > 0 (no dx*dy factor, as the reference).  Returns the
 * triplet count, or a negative code (also when more than buffer_size triplets
 * are needed; the reference asserts).  B1 rows are the footprint of
 * b1_forward_projection: the reference's matrix kernel averages over
 * [y, y + vblur] (A_B1.cu:402-411), which returns bin u-1's value in row u;
 * its projector and its quadrature generator python/matrix_gen/
 * A_B1_ref_project.py agree with the rows returned here.  Addition: with
 * spmat == NULL and buffer_size == 0 (this and the exporters below) the call
 * only counts, and returns the number of triplets the matrix needs. */
int pixel_matrix_generation(int np, int nu, int nx, int ny, double dx, double dy, double du, double lsd,
                            double lso, double *spmat, int buffer_size);
int bilinear_matrix_generation(int np, int nu, int nx, int ny, double dx, double dy, double du, double lsd,
                               double lso, double *spmat, int buffer_size);

/* The hexagonal exporters python/matrix_gen/A_H0_project.py:17-19 and
 * A_Hl_project.py:17-19 bind (h0_matrix_generation; hex_matrix_generation for
 * Hl).  The reference's own versions are commented out (A_H0.cu:510-553,
 * A_Hl.cu:521-565; their kernels A_H0.cu:360, A_Hl.cu:370-451 are never
 * launched).  nl rows x nc columns of lattice constant dt, np views over
 * [0, 2 pi (np-1)/np]; COO triplets (p*nu + u, il*nc + ic, conv) in the order
 * view, bin, il, ic, kept where conv > 0, conv being the footprint of
 * h0_/hl_forward_projection without the dx dy factor (H0 with the 1/3 and
 * vertex-bound fixes above).  Returns the triplet count or a negative code
 * (also when more than buffer_size triplets are needed). */
int h0_matrix_generation(int np, int nu, int nl, int nc, double du, double dt, double lsd, double lso,
                         double *spmat, int buffer_size);
int hl_matrix_generation(int np, int nu, int nl, int nc, double du, double dt, double lsd, double lso,
                         double *spmat, int buffer_size);
int hex_matrix_generation(int np, int nu, int nl, int nc, double du, double dt, double lsd, double lso,
                          double *spmat, int buffer_size);

/* layers/A_B1_3D.hh:24-26 / A_B1_3D.cu:701 (b1_3d_init), :709 (forward),
 * :744 (backward).  Cone beam, in-plane B1 footprint times an axial
 * trapezoid; volume x copied raw as [iz][iy][ix] (A_B1_3D.cu:716),
 * projections b = [p][v][u] (:731).  Forward returns the reference's
 * evaluation count; a value >= 2^63 is a negative SGDCT_E* code.  The
 * reference's back-projector returns zeros (A_B1_3D.cu:408, 436 multiply by
 * the zeroed output); this one is the transpose of the forward and returns
 * SGDCT_OK or a negative code.  Any nz and nv (columns of more than 512
 * voxels and detectors of more than 512 rows are processed in chunks). */
sgdct_projector *b1_3d_init(int nx, int ny, int nz, int np, int nu, int nv, double dx, double dy,
                            double dz, double du, double dv, double lsd, double lso,
                            const double *angles);
unsigned long long b1_3d_forward_projection(double *b, double *x, sgdct_projector *h);
int b1_3d_backward_projection(double *b, double *x, sgdct_projector *h);

/* layers/SF_projector.hh / SF_projector.cu:81-400 (class SF; the reference has
 * no C entry points for it): cone-beam separable footprints, trap_v = 0 SF-TR
 * (rectangle in v), 1 SF-TT (trapezoid in v).  Same layouts as b1_3d_*;
 * detector cells centred on integer coordinates (SF_projector.cu:209-216).
 * Fixed against the reference: the value carries the voxel and the amplitude
 * lsd^2 dx dy dz / (du dv r^2) (its forward adds only the footprint product,
 * :230-231, and :364 has dy dz), SF-TT's v area is 2 / ((v3-v0) + (v2-v1))
 * (:195 misplaces the parenthesis), r is measured from the voxel centre (:103
 * uses nx / 2 in integers), cells off the detector are dropped (:228 wraps
 * into the next row), and backward is the transpose of forward (:394 scales
 * by lsd dx / np).  Both return SGDCT_OK or a negative code. */
sgdct_projector *sf3d_init(int nx, int ny, int nz, int np, int nu, int nv, double dx, double dy, double dz,
                           double du, double dv, double lsd, double lso, const double *angles, int trap_v);
int sf3d_forward_projection(double *b, double *x, sgdct_projector *h);
int sf3d_backward_projection(double *b, double *x, sgdct_projector *h);

/* ---- additions ------------------------------------------------------------ */

/* SF_init with per-view source offsets, as b0_init / b1_init take them
 * (layers/A_B1.hh:24; utils/geo.cu:243-251: the source of view p sits at
 * R(angle_p)(-lso + drho[p], 0, dz[p])), for fan slices rebinned from a
 * helical scan; the reference's SF_init (A_SF.cu:396) has no offsets.  Either
 * pointer may be NULL; dz does not change a 2-D fan projection. */
sgdct_projector *sgdct_sf_init_offsets(int nx, int ny, int np, int nu, double dx, double dy, double du,
                                       double lsd, double lso, const double *angles, const double *dz,
                                       const double *drho);

/* Release a handle from any *_init (the reference leaks them). */
void sgdct_destroy(sgdct_projector *h);

/* Image (ny, nx) FP32 row-major [iy][ix] and sinogram (np, nu) FP32 [p][u]
 * (cone beam: volume (nz, ny, nx) and projections (np, nv, nu)), both device
 * pointers; `batch` images / sinograms back to back; `stream` is
 * a hipStream_t (NULL = default stream).  Asynchronous on `stream`.  Cone-beam
 * handles, and fan-beam handles on launches they split (below), keep per-handle
 * workspace: calls on one handle must not overlap (issue them on one stream).
 * A cone-beam A that takes more than one pass of views (a pass holds up to
 * 2 GiB of row profiles) also runs every other pass on a stream of the
 * handle's own, joined back to `stream` by events before the call returns;
 * its second 2 GiB buffer is allocated with the rest of the workspace.
 * Returns SGDCT_OK or a negative code. */
int sgdct_forward_dev(sgdct_projector *h, const float *img, float *sino, int batch, void *stream);
int sgdct_backward_dev(sgdct_projector *h, const float *sino, float *img, int batch, void *stream);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  sgdct_forward_dev (direction 0: in = images, out = sinograms)
 * or sgdct_backward_dev (direction 1: in = sinograms, out = images) on a
 * fan-beam handle (B0, B1, SF, H0, Hl), with the FP64 partial sums of split
 * launches in a static device array of 32 MiB per GPU (allocated when the
 * library's code is loaded on that GPU, held until the process ends) instead
 * of the handle's workspace.  The call allocates nothing; its first call on a
 * GPU looks up the array's address there.  The handle's device must be
 * current.  A split chosen automatically is lowered, when one image's partial
 * sums do not fit in 32 MiB, to the largest that fits (results agree with
 * sgdct_*_dev to FP64 rounding of the partial sums; bit for bit at the same
 * split); a split fixed by sgdct_split_config that does not fit is
 * SGDCT_EINVAL.  Cone-beam handles are SGDCT_EUNSUP.  Every call on a GPU
 * shares that GPU's array (each launch writes every partial sum it reads, so
 * nothing needs clearing): calls that overlap on one GPU, from several
 * streams or threads, race. */
int sgdct_project_dev_noalloc(sgdct_projector *h, int direction, const float *in, float *out, int batch,
                              void *stream);

/* Fan-beam launch shape (no reference counterpart).  A splits every bin's
 * pixel walk into `fwd_split` line segments (on hexagonal handles: segments of
 * the lattice rows, or of the vertical lattice lines a wave walks when its
 * strips are steep) and A^T the views into `bwd_split` groups, each writing
 * FP64 partial sums that a second kernel adds
 * in a fixed order (results are deterministic, independent of the batch size,
 * and agree across splits to FP64 rounding).  0 = automatic: chosen at
 * creation from one image's launch size (a launch of few waves splits, up to
 * 24 ways; hexagonal A stays unsplit); 1..32 = fixed.  The partial-sum workspace (2 images per launch,
 * or up to 16 while it stays within 64 MiB; larger batches take several
 * launches) is allocated at creation and by this call, never inside a
 * projection.  Returns SGDCT_OK, SGDCT_EINVAL or SGDCT_ENOMEM. */
int sgdct_split_config(sgdct_projector *h, int fwd_split, int bwd_split);

/* Shape query: out[0..3] = nx, ny, np, nu (hexagonal: nc, nl, np, nu);
 * returns the basis id (0 B0, 1 B1, 2 SF, 3 H0, 4 Hl) or a negative code. */
int sgdct_shape(const sgdct_projector *h, int *out);

/* Shape query for any handle: out[0..5] = nx, ny, nz, np, nu, nv (nz = nv = 1
 * for the fan-beam bases); returns the basis id (5 = cone-beam B1, 6 = cone-beam
 * separable footprints). */
int sgdct_shape3(const sgdct_projector *h, int *out);

/* Footprint evaluations of one A under the reference's corner rule (the
 * count layers/A_B1.cu:643 returns), as 64-bit; negative on error. */
long long sgdct_eval_count(sgdct_projector *h);

/* Iterative-reconstruction kernels (device pointers, FP32, asynchronous on
 * `stream`).  No reference counterpart: optimizer/asd_pocs.cu is empty; these
 * implement the per-pixel steps of SIRT and ASD-POCS (Sidky & Pan 2008).
 *   tv_gradient:   g = d/dx sum sqrt(eps + dx^2 + dy^2), backward differences
 *   sirt_residual: r = w * (b - ax), w broadcast over n / per batch entries
 *   sirt_update:   x = x + lam * v * c, then max(x, 0) if positivity
 *   axpy:          x = x + a * g */
int sgdct_tv_gradient(const float *x, float *g, int nx, int ny, int batch, float eps, void *stream);
int sgdct_sirt_residual(const float *b, const float *ax, const float *w, float *r, long long n,
                        long long per, void *stream);
int sgdct_sirt_update(float *x, const float *c, const float *v, float lam, int positivity, long long n,
                      long long per, void *stream);
int sgdct_axpy(float *x, const float *g, float a, long long n, void *stream);
/* x[k] += a[k / per] g[k], a a device array of n / per coefficients */
int sgdct_axpy_batch(float *x, const float *g, const float *a, long long n, long long per, void *stream);
/* `nsteps` steps of ASD-POCS's normalised TV descent on a batch of images,
 * with no host synchronisation between them: g = tv_gradient(x), then
 * x[b] -= step[b] * g[b] / ||g[b]|| for each image with a nonzero gradient
 * (||g[b]|| summed in FP64 in a fixed order; deterministic).  step: `batch`
 * FP64 values on the device (0 leaves an image unchanged); g: device space
 * for the gradients (nx * ny * batch floats); work: 64 * batch doubles. */
int sgdct_tv_descent(float *x, float *g, double *work, const double *step, int nx, int ny, int batch,
                     float eps, int nsteps, void *stream);

/* ---- batched slices (no reference counterpart: the reference projects one
 * slice per call and launches one kernel per view, layers/A_B1.cu:512-541) --
 * Many B0 / B1 / SF projectors of uneven size, view count and basis run as
 * one persistent launch per direction over a device work queue (A: 64-bin
 * waves of one view, A^T: 16x16 pixel tiles), slices ordered by decreasing
 * per-item cost.  Results equal the per-projector entry points to one FP32
 * rounding per element (those may split a walk or the views into FP64
 * partial sums, associated differently), and bit for bit in the per-slice
 * mode of sgdct_batch_config.  Images / sinograms of all slices are packed back to back in one
 * device buffer each, at the element offsets sgdct_batch_layout returns
 * (nslice + 1 entries each).  The projectors must outlive the batch; calls on
 * one batch must not overlap (they share the queue counters). */
typedef struct sgdct_batch sgdct_batch;
sgdct_batch *sgdct_batch_create(sgdct_projector *const *slices, int nslice);
void sgdct_batch_destroy(sgdct_batch *b);
int sgdct_batch_layout(const sgdct_batch *b, long long *img_off, long long *sino_off);
/* Launch shape: persistent blocks per CU for A and A^T (1..32, or 0 for one
 * launch per slice through the slices' own projector kernels and workspaces)
 * and slice order of the persistent queue (0: decreasing per-item cost, 1: as
 * given); set by the autotuner.  Defaults: 8, 8, 0. */
int sgdct_batch_config(sgdct_batch *b, int fwd_blocks_per_cu, int bwd_blocks_per_cu, int order);
int sgdct_batch_forward(sgdct_batch *b, const float *img, float *sino, void *stream);
int sgdct_batch_backward(sgdct_batch *b, const float *sino, float *img, void *stream);
/* Same as forward (direction 0) / backward (1), through a hipGraph captured on
 * `stream` (non-default) at the first call for these buffers and replayed
 * afterwards; the graph clears its queue counters with a kernel, not a memset. */
int sgdct_batch_graph_launch(sgdct_batch *b, int direction, const float *in, float *out, void *stream);

/* ---- warm-start network kernels (no reference counterpart; the reference's
 * network code is a cuBLAS MLP, network.cu and layers/linear.cu).  NHWC
 * bfloat16 activations (passed as void*), FP32 gradients of activations,
 * weights bfloat16 W[cout][9 cin] (k = (3 ky + kx) cin + c), FP32 master
 * weights.  All asynchronous on `stream`; workspaces are caller-owned device
 * FP32 buffers of at least the size the matching *_workspace call returns. */
/* out = relu?(conv3x3(in, w) + bias), zero padding; bf16 out, or FP32 when
 * out_f32 (the input-gradient pass: in = bf16 output gradient, w = flipped). */
int sgdct_nn_conv3x3(const void *in, const void *w, const float *bias, void *out, int B, int H, int W, int cin,
                     int cout, int relu, int out_f32, void *stream);
/* sgdct_nn_conv3x3 with two fusions.  in2 != NULL: the input is the decoder's
 * concatenation [upsample2(in), in2], read in place: in holds c0 channels at
 * H/2 x W/2 (nearest-neighbour upsampled), in2 the other cin - c0 channels at
 * H x W (c0 a multiple of 8, H and W even).  mask != NULL ([B][H][W][cout]
 * bf16): each output value is kept where mask > 0 and zeroed elsewhere (the
 * ReLU mask of an activation gradient, applied before the bf16 / FP32 store). */
int sgdct_nn_conv3x3_ex(const void *in, const void *in2, int c0, const void *w, const float *bias, void *out,
                        const void *mask, int B, int H, int W, int cin, int cout, int relu, int out_f32,
                        void *stream);
long long sgdct_nn_wgrad_workspace(int B, int H, int W, int cin, int cout);
/* dw[co][k] = sum over pixels of dy[m][co] x[m + tap(k)][c(k)] */
int sgdct_nn_wgrad(const void *x, const void *dy, float *dw, float *workspace, long long workspace_floats, int B,
                   int H, int W, int cin, int cout, void *stream);
/* sgdct_nn_wgrad of the concatenated input [upsample2(x), x2] (as
 * sgdct_nn_conv3x3_ex; x2 = NULL is sgdct_nn_wgrad), and with db != NULL also
 * the bias gradient db[co] = sum over pixels of dy[m][co] (FP32 sums). */
int sgdct_nn_wgrad_ex(const void *x, const void *x2, int c0, const void *dy, float *dw, float *db, float *workspace,
                      long long workspace_floats, int B, int H, int W, int cin, int cout, void *stream);
long long sgdct_nn_colsum_workspace(long long M, int C);
int sgdct_nn_colsum(const void *a, float *out, float *workspace, long long workspace_floats, long long M, int C,
                    void *stream);
int sgdct_nn_pad_input(const float *x, void *out8, long long npix, void *stream);
/* The gradients below that end in a ReLU activation's mask can be written
 * masked and in bf16 directly: given out (bf16) and mask (the activation,
 * bf16), out = bf16(value * (mask > 0)) replaces the FP32 result. */
/* dst (B, H, W, C) = 2x2 average of src (B, 2H, 2W, C); back: gsrc += dv / 4
 * (with out: out = masked gsrc + dv / 4, gsrc unchanged) */
int sgdct_nn_pool2(const void *src, void *dst, int B, int H, int W, int C, void *stream);
int sgdct_nn_pool2_back(const float *dv, float *gsrc, void *out, const void *mask, int B, int H, int W, int C,
                        void *stream);
/* dst (B, H, W, C0 + C1) = [up2(low (B, H/2, W/2, C0)), skip (B, H, W, C1)];
 * back: glow = sum over each 2x2 block of dv[..., :C0] (with glow_out: masked
 * into glow_out), gskip = dv[..., C0:] (+= with skip_accumulate) */
int sgdct_nn_upcat(const void *low, const void *skip, void *dst, int B, int H, int W, int C0, int C1, void *stream);
int sgdct_nn_upcat_back(const float *dv, float *glow, void *glow_out, const void *glow_mask, float *gskip,
                        int skip_accumulate, int B, int H, int W, int C0, int C1, void *stream);
/* y = x + d0 . w + b (1x1 output head); back: dy = 2 (y - target) / M,
 * gd0 = dy w (with gd0_out: masked into gd0_out), gw[0:C] = sum dy d0,
 * gw[C] = sum dy */
int sgdct_nn_head(const void *d0, const float *w, const float *b, const float *x, float *y, long long M, int C,
                  void *stream);
long long sgdct_nn_head_workspace(long long M, int C);
int sgdct_nn_head_back(const float *y, const float *target, const float *w, const void *d0, float *dy, float *gd0,
                       void *gd0_out, const void *mask, float *gw, float *workspace, long long workspace_floats,
                       long long M, int C, void *stream);
/* out = bf16(g * (act > 0)) */
int sgdct_nn_relu_mask(const float *g, const void *act, void *out, long long n, void *stream);
/* wt[c][t][o] = w[o][8 - t][c] */
int sgdct_nn_flip_weights(const void *w, void *wt, int cin, int cout, void *stream);
int sgdct_nn_adam(float *w32, void *w16, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                  float eps, int step, void *stream);
int sgdct_nn_fill(float *p, long long n, float value, void *stream);
int sgdct_nn_to_bf16(const float *src, void *dst, long long n, void *stream);

/* Human-readable description of the calling thread's last error. */
const char *sgdct_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SGDCT_H */
