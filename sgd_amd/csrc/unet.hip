// unet.hip -- gfx950 training and inference kernels of the warm-start U-Net.
//
// No reference counterpart (the reference's network code is a cuBLAS MLP,
// network.cu / layers/linear.cu; its Layer::forward/back_prop pattern,
// layers/layer.hh:15-16, is the closest precedent).  Activations are NHWC
// bfloat16, gradients of activations FP32, weights bfloat16 W[cout][9 cin]
// (tap-major: k = (3 ky + kx) cin + c) with an FP32 master copy.
//
//   conv3x3_kernel   3x3 zero-padded convolution as an implicit GEMM on the
//                    matrix cores: v_mfma_f32_16x16x32_bf16, FP32 accumulate.
//                    Block = 4 waves, each a 32-pixel x 32-channel tile
//                    (2 x 2 MFMA tiles); operands are gathered straight from
//                    HBM/L2 as 16-byte vectors (8 channels of one tap).
//                    Bias + ReLU fused; stores bf16, or FP32 for the
//                    input-gradient pass (which runs the same kernel on the
//                    output gradient with the flipped, transposed weights).
//   wgrad_kernel     dW[co][k] = sum_m dY[m][co] X[m + tap(k)][c(k)]: the
//                    pixel sum is the GEMM's K; tiles of dY and X are staged
//                    in LDS and read transposed; each block reduces a
//                    contiguous pixel range into an FP32 partial, and a
//                    second kernel sums the partials in a fixed order.
//   small kernels    input padding, 2x2 average pool, nearest upsample +
//                    concatenation and their adjoints, the 1x1 output head,
//                    MSE gradient, ReLU-mask + bf16 conversion, weight
//                    transpose, Adam.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sgdct {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) i16x4 lds_i16x4;

struct ConvArgs {
  const bf16* in;      // [B][H][W][cin] (with in2: [B][H/2][W/2][c0])
  const bf16* w;       // [cout][9 cin]
  const float* bias;   // [cout] or null
  void* out;           // [B][H][W][cout], bf16 or FP32
  int B, H, W, cin, cout;
  int relu, out_f32;
  const bf16* mask;    // [B][H][W][cout] or null: out = value * (mask > 0)
  const bf16* in2;     // [B][H][W][cin - c0] or null (see conv_in)
  int c0;
};

// Address of the 8 input channels c .. c + 7 at pixel (b, y, x).  With in2 the
// input is the decoder's concatenation, read in place: channels [0, c0) are
// the half-resolution tensor `in` upsampled by nearest neighbour, channels
// [c0, cin) the full-resolution skip tensor in2 (c0 % 8 == 0, so a group of
// 8 never straddles the two).
__device__ __forceinline__ const bf16* conv_in(const bf16* in, const bf16* in2, int c0, int H, int W, int cin, int b,
                                               int y, int x, int c) {
  if (!in2) return in + (((size_t)b * H + y) * W + x) * cin + c;
  if (c < c0) return in + (((size_t)b * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * c0 + c;
  return in2 + (((size_t)b * H + y) * W + x) * (cin - c0) + (c - c0);
}

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (bf16)0.0f;
  return z;
}

// One 32-pixel x 32-channel wave tile per (wm, wn); block tile WM*32 x WN*32.
template <int WM, int WN>
__global__ __launch_bounds__(256) void conv3x3_kernel(ConvArgs a) {
  static_assert(WM * WN == 4, "four waves per block");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wm = wv % WM, wn = wv / WM;
  const long long M = (long long)a.B * a.H * a.W;
  const long long m_base = (long long)blockIdx.x * (WM * 32) + wm * 32;
  const int n_base = blockIdx.y * (WN * 32) + wn * 32;
  const int K = 9 * a.cin;
  const int r16 = lane & 15, g = lane >> 4;

  // pixel coordinates of the two A rows this lane loads
  int py[2], px[2], pb[2];
  bool pv[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const long long m = m_base + 16 * i + r16;
    pv[i] = m < M;
    const long long mm = pv[i] ? m : 0;
    px[i] = (int)(mm % a.W);
    const long long t = mm / a.W;
    py[i] = (int)(t % a.H);
    pb[i] = (int)(t / a.H);
  }
  int co[2];
  bool cv[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    co[j] = n_base + 16 * j + r16;
    cv[j] = co[j] < a.cout;
  }

  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int k0 = 0; k0 < K; k0 += 32) {
    const int k = k0 + 8 * g;  // this lane's 8 consecutive k: one tap, 8 channels
    const bool kv = k < K;
    const int t = kv ? k / a.cin : 0, c = kv ? k - t * a.cin : 0;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    bf16x8 av[2], bv[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int yy = py[i] + dy, xx = px[i] + dx;
      if (kv && pv[i] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
        av[i] = *reinterpret_cast<const bf16x8*>(conv_in(a.in, a.in2, a.c0, a.H, a.W, a.cin, pb[i], yy, xx, c));
      else
        av[i] = zero8();
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (kv && cv[j])
        bv[j] = *reinterpret_cast<const bf16x8*>(a.w + (size_t)co[j] * K + k);
      else
        bv[j] = zero8();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
  }

  // D layout: lane holds rows 4 g + r (r = 0..3) of column r16
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if (!cv[j]) continue;
    const float bias = a.bias ? a.bias[co[j]] : 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long m = m_base + 16 * i + 4 * g + r;
        if (m >= M) continue;
        float v = acc[i][j][r] + bias;
        if (a.relu) v = fmaxf(v, 0.0f);
        if (a.mask && !((float)a.mask[m * a.cout + co[j]] > 0.0f)) v = 0.0f;
        if (a.out_f32) static_cast<float*>(a.out)[m * a.cout + co[j]] = v;
        else static_cast<bf16*>(a.out)[m * a.cout + co[j]] = (bf16)v;
      }
  }
}

// ---------------------------------------------------------------------------
// LDS-tiled implicit GEMM for the same convolution (cout a multiple of 32).
// Block tile BM pixels x BN output channels, K in steps of 32 (k = tap-major,
// each 8-element group one tap and 8 channels since cin % 8 == 0).  Every
// K step the block stages A (BM x 32) and B (BN x 32) in LDS with 80-byte
// rows (16 bytes of padding: the 16-byte fragment reads of 8 consecutive rows
// then fall in distinct banks), double-buffered: the next step's global loads
// are in flight while the waves run their MFMAs on the current one, and one
// barrier per step separates the two buffers' uses.  Waves WM x WN, each
// (BM / WM) x (BN / WN) of MFMA 16x16x32 tiles.
// ---------------------------------------------------------------------------
constexpr int kLdsRow = 40;  // bf16 per LDS row: 32 + 8 padding

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(256) void conv3x3_lds_kernel(ConvArgs a) {
  static_assert(WM * WN == 4, "four waves per block");
  constexpr int TM = BM / WM / 16, TN = BN / WN / 16;  // MFMA tiles per wave
  constexpr int AC = BM * 4 / 256;                     // A chunks (16 B) per thread
  constexpr int BC = (BN * 4 + 255) / 256;             // B chunks per thread (some idle if BN = 32)
  __shared__ __attribute__((aligned(16))) bf16 As[2][BM * kLdsRow];
  __shared__ __attribute__((aligned(16))) bf16 Bs[2][BN * kLdsRow];

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = wv % WM, wn = wv / WM;
  const long long M = (long long)a.B * a.H * a.W;
  const long long m_base = (long long)blockIdx.x * BM;
  const int n_base = blockIdx.y * BN;
  const int K = 9 * a.cin, nk = (K + 31) / 32;
  const int r16 = lane & 15, g = lane >> 4;

  // this thread's A chunks: row = q / 4, k group = q % 4, q = tid + 256 i
  int ay[AC], ax[AC], ab[AC];
  bool av_ok[AC];
#pragma unroll
  for (int i = 0; i < AC; ++i) {
    const int q = tid + 256 * i;
    const long long m = m_base + q / 4;
    av_ok[i] = m < M;
    const long long mm = av_ok[i] ? m : 0;
    ax[i] = (int)(mm % a.W);
    const long long t = mm / a.W;
    ay[i] = (int)(t % a.H);
    ab[i] = (int)(t / a.H);
  }

  bf16x8 ra[AC], rb[BC];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < AC; ++i) {
      const int q = tid + 256 * i;
      const int k = k0 + 8 * (q & 3);
      bf16x8 v = zero8();
      if (av_ok[i] && k < K) {
        const int t = k / a.cin, c = k - t * a.cin;
        const int yy = ay[i] + t / 3 - 1, xx = ax[i] + t % 3 - 1;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
          v = *reinterpret_cast<const bf16x8*>(conv_in(a.in, a.in2, a.c0, a.H, a.W, a.cin, ab[i], yy, xx, c));
      }
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < BC; ++i) {
      const int q = tid + 256 * i;
      const int n = q / 4, k = k0 + 8 * (q & 3);
      bf16x8 v = zero8();
      if (q < BN * 4 && n_base + n < a.cout && k < K)
        v = *reinterpret_cast<const bf16x8*>(a.w + (size_t)(n_base + n) * K + k);
      rb[i] = v;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < AC; ++i) {
      const int q = tid + 256 * i;
      *reinterpret_cast<bf16x8*>(&As[buf][(q / 4) * kLdsRow + 8 * (q & 3)]) = ra[i];
    }
#pragma unroll
    for (int i = 0; i < BC; ++i) {
      const int q = tid + 256 * i;
      if (q < BN * 4) *reinterpret_cast<bf16x8*>(&Bs[buf][(q / 4) * kLdsRow + 8 * (q & 3)]) = rb[i];
    }
  };

  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * 32);
    bf16x8 fa[TM], fb[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
      fa[i] = *reinterpret_cast<const bf16x8*>(&As[buf][(wm * (BM / WM) + 16 * i + r16) * kLdsRow + 8 * g]);
#pragma unroll
    for (int j = 0; j < TN; ++j)
      fb[j] = *reinterpret_cast<const bf16x8*>(&Bs[buf][(wn * (BN / WN) + 16 * j + r16) * kLdsRow + 8 * g]);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    if (kt + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }

  // D layout: lane holds rows 4 g + r (r = 0..3) of column r16
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int co = n_base + wn * (BN / WN) + 16 * j + r16;
    if (co >= a.cout) continue;
    const float bias = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long long m = m_base + wm * (BM / WM) + 16 * i + 4 * g + r;
        if (m >= M) continue;
        float v = acc[i][j][r] + bias;
        if (a.relu) v = fmaxf(v, 0.0f);
        if (a.mask && !((float)a.mask[m * a.cout + co] > 0.0f)) v = 0.0f;
        if (a.out_f32) static_cast<float*>(a.out)[m * a.cout + co] = v;
        else static_cast<bf16*>(a.out)[m * a.cout + co] = (bf16)v;
      }
  }
}

// ---------------------------------------------------------------------------
// Halo-tiled 3x3 convolution, the same implicit GEMM with its operands staged
// once per block and input-channel chunk: the block's output tile is TH x 16
// pixels x BN channels, and for each chunk of CC input channels it stages in
// LDS the (TH + 2) x 18 pixel halo of the tile (80-byte pixel rows for
// CC = 32) and the chunk's BN x 9 x CC weights (rows of 9 CC + 8 bf16), from
// which every tap's A and B fragments are read; the next chunk's operands are
// loaded into registers meanwhile.  Nine taps thus reuse one staged slab
// instead of re-gathering the tile's pixels per tap.  CC = 32 (cin % 32 == 0):
// a K step is one tap x 32 channels; CC = 8 (cin == 8): a K step is four
// taps x 8 channels (lane group g takes tap 4 s + g).  Wave (wm, wn)
// computes TM tile rows x TN 16-channel tiles.
// ---------------------------------------------------------------------------
template <int TH, int BN, int WM, int WN, int CC>
__global__ __launch_bounds__(256) void conv3x3_halo_kernel(ConvArgs a) {
  static_assert(WM * WN == 4 && TH % WM == 0, "four waves, whole tile rows per wave");
  constexpr int TW = 16, HH = TH + 2, HWD = TW + 2;
  constexpr int PS = CC == 32 ? 40 : CC;          // LDS pixel stride (bf16)
  constexpr int WR = 9 * CC + 8;                  // LDS weight row (bf16)
  constexpr int TM = TH / WM, TN = BN / WN / 16;  // MFMA tiles per wave
  constexpr int GP = CC / 8;                      // 16-byte groups per halo pixel / weight tap
  constexpr int NG = HH * HWD * GP;               // groups per halo
  constexpr int NW = BN * 9 * GP;                 // groups per weight slab
  constexpr int HC = (NG + 255) / 256, WC = (NW + 255) / 256;  // groups per thread
  constexpr int TPS = 32 / CC;                    // taps per K step
  constexpr int NKS = (9 + TPS - 1) / TPS;        // K steps per chunk
  constexpr int ES = BN + 4;                      // epilogue tile row (floats)
  constexpr int NSTAGE = HH * HWD * PS + BN * WR, NEPI = TH * TW * ES * 2;  // bf16 units
  __shared__ __attribute__((aligned(16))) bf16 smem[NSTAGE > NEPI ? NSTAGE : NEPI];
  bf16* const halo = smem;
  bf16* const wts = smem + HH * HWD * PS;

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = wv % WM, wn = wv / WM;
  const int r16 = lane & 15, g = lane >> 4;
  const int tiles_x = (a.W + TW - 1) / TW, tiles_y = (a.H + TH - 1) / TH;
  int t = blockIdx.x;
  const int tx = t % tiles_x;
  t /= tiles_x;
  const int ty = t % tiles_y, b = t / tiles_y;
  const int x0 = tx * TW, y0 = ty * TH;
  const int nb = blockIdx.y * BN;
  const int K = 9 * a.cin, nchunk = a.cin / CC;

  bf16x8 rh[HC], rw[WC];
  auto load = [&](int cc) {
#pragma unroll
    for (int i = 0; i < HC; ++i) {
      const int q = tid + 256 * i, p = q / GP, cg = (q % GP) * 8;
      const int hy = p / HWD, hx = p - hy * HWD;
      const int iy = y0 + hy - 1, ix = x0 + hx - 1;
      bf16x8 v = zero8();
      if (q < NG && iy >= 0 && iy < a.H && ix >= 0 && ix < a.W)
        v = *reinterpret_cast<const bf16x8*>(conv_in(a.in, a.in2, a.c0, a.H, a.W, a.cin, b, iy, ix, cc * CC + cg));
      rh[i] = v;
    }
#pragma unroll
    for (int i = 0; i < WC; ++i) {
      const int q = tid + 256 * i, n = q / (9 * GP), rem = q - n * (9 * GP);
      const int tp = rem / GP, cg = (rem - tp * GP) * 8;
      bf16x8 v = zero8();
      if (q < NW && nb + n < a.cout)
        v = *reinterpret_cast<const bf16x8*>(a.w + (size_t)(nb + n) * K + tp * a.cin + cc * CC + cg);
      rw[i] = v;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < HC; ++i) {
      const int q = tid + 256 * i, p = q / GP, cg = (q % GP) * 8;
      if (q < NG) *reinterpret_cast<bf16x8*>(&halo[p * PS + cg]) = rh[i];
    }
#pragma unroll
    for (int i = 0; i < WC; ++i) {
      const int q = tid + 256 * i, n = q / (9 * GP), rem = q - n * (9 * GP);
      if (q < NW) *reinterpret_cast<bf16x8*>(&wts[n * WR + rem * 8]) = rw[i];
    }
  };

  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  // per-lane fragment bases: A at tile row wm TM, pixel r16, channels
  // (8 g) % CC; B at weight row wn BN / WN + r16
  const bf16* const abase = &halo[((wm * TM) * HWD + r16) * PS + (8 * g) % CC];
  const bf16* const bbase = &wts[(wn * (BN / WN) + r16) * WR + (8 * g) % CC];
  load(0);
  for (int cc = 0; cc < nchunk; ++cc) {
    if (cc > 0) __syncthreads();  // the previous chunk's reads come first
    store();
    __syncthreads();
    if (cc + 1 < nchunk) load(cc + 1);
#pragma unroll 1
    for (int ks = 0; ks < NKS; ++ks) {
      // this lane's tap in K step ks and its halo offsets (0..2)
      const int tg = ks * TPS + (8 * g) / CC;
      const int dy = tg / 3, dx = tg - 3 * dy;
      const bool tv = tg < 9;
      const bf16* const ap = abase + (dy * HWD + dx) * PS;
      const bf16* const bp = bbase + tg * CC;
      bf16x8 fa[TM], fb[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) fa[i] = tv ? *reinterpret_cast<const bf16x8*>(ap + i * HWD * PS) : zero8();
#pragma unroll
      for (int j = 0; j < TN; ++j) fb[j] = tv ? *reinterpret_cast<const bf16x8*>(bp + 16 * j * WR) : zero8();
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
  }

  // Epilogue through LDS: the block's FP32 tile [pixel][channel] (row stride
  // BN + 4 floats) replaces the staged operands, then each thread finishes 8
  // channels of one pixel with 16-byte loads (mask) and stores.
  constexpr int EG = TH * TW * BN / 8;  // epilogue groups
  float* const etile = reinterpret_cast<float*>(smem);
  __syncthreads();
  // D layout: lane holds rows 4 g + r (pixels of tile row wm TM + i) of column r16
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        etile[((wm * TM + i) * TW + 4 * g + r) * ES + wn * (BN / WN) + 16 * j + r16] = acc[i][j][r];
  __syncthreads();
  for (int q = tid; q < EG; q += 256) {
    const int p = q / (BN / 8), c8 = (q - p * (BN / 8)) * 8;
    const int oy = y0 + p / TW, ox = x0 + p % TW, co = nb + c8;
    if (oy >= a.H || ox >= a.W || co >= a.cout) continue;
    const long long o = (((long long)b * a.H + oy) * a.W + ox) * a.cout + co;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      v[e] = etile[p * ES + c8 + e] + (a.bias ? a.bias[co + e] : 0.0f);
      if (a.relu) v[e] = fmaxf(v[e], 0.0f);
    }
    if (a.mask) {
      const i16x8 mk = __builtin_bit_cast(i16x8, *reinterpret_cast<const bf16x8*>(a.mask + o));
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (!(__uint_as_float((unsigned)(unsigned short)mk[e] << 16) > 0.0f)) v[e] = 0.0f;
    }
    if (a.out_f32) {
      float* const op = static_cast<float*>(a.out) + o;
      *reinterpret_cast<f32x4*>(op) = f32x4{v[0], v[1], v[2], v[3]};
      *reinterpret_cast<f32x4*>(op + 4) = f32x4{v[4], v[5], v[6], v[7]};
    } else {
      bf16x8 h;
#pragma unroll
      for (int e = 0; e < 8; ++e) h[e] = (bf16)v[e];
      *reinterpret_cast<bf16x8*>(static_cast<bf16*>(a.out) + o) = h;
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradient dW[co][k] = sum_m dY[m][co] X[m + tap(k)][c(k)]: a GEMM
// whose reduction runs over pixels m.  Block tile CO_T (co) x KT (k) with
// KT = 256 / CO_T * 16 ... 4 waves, each a 16-row co strip x 64 k (4 MFMA
// tiles).  The block walks its pixel range in chunks of 64.  The chunk's dY
// and X tiles are staged in LDS in their natural pixel-major layout
// ([pixel][channel], one 16-byte write per 8 channels), and gfx950's
// transposing LDS read (ds_read_b64_tr_b16) hands each lane the 8
// consecutive pixels of one channel that an MFMA fragment needs, as two
// 8-byte reads.  Within each 16-lane group that read gives lane i, value k,
// the 16-bit element i % 4 of the 8 bytes addressed by lane 4k + i / 4; so
// lane j addresses pixel row p0 + j / 4, channels c0 + 4 (j % 4) .. + 3.
// Row strides of 80 / 48 / 144 bf16 (160 / 96 / 288 bytes, 32 or 96 mod 128)
// put the 4 rows a 16-lane group reads in distinct banks.  Result ->
// partial[blockIdx.z] (pixel-range splits, added in a fixed order by
// reduce_partials_kernel).
// ---------------------------------------------------------------------------
struct WgradArgs {
  const bf16* x;    // [B][H][W][cin] (with x2: the concatenation, as ConvArgs::in2)
  const bf16* dy;   // [B][H][W][cout]
  float* partial;   // [nsplit][cout][9 cin]
  int B, H, W, cin, cout;
  long long pix_per_split;  // multiple of 64
  const bf16* x2;
  int c0;
  float* bias_partial;      // [nsplit][cout] column sums of dy, or null
};


// An MFMA fragment of 8 consecutive pixels of one channel from a
// pixel-major LDS tile: the transposing reads of pixel rows [p, p + 4) and
// [p + 4, p + 8) (p, c as addressed by this lane, see above).
__device__ __forceinline__ bf16x8 tr_fragment(const bf16* tile, int row, int col, int stride) {
  const lds_i16x4* p0 = (const lds_i16x4*)((__attribute__((address_space(3))) char*)tile +
                                           2 * (row * stride + col));
  const lds_i16x4* p1 = (const lds_i16x4*)((__attribute__((address_space(3))) char*)tile +
                                           2 * ((row + 4) * stride + col));
  const i16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16(const_cast<lds_i16x4*>(p0));
  const i16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16(const_cast<lds_i16x4*>(p1));
  const i16x8 v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, v);
}

template <int CO_T>
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a) {
  static_assert(CO_T == 64 || CO_T == 32, "co tile 64 or 32");
  constexpr int WK = 64 / CO_T * 1;       // waves along k: 1 (CO_T 64) or 2 (CO_T 32)
  constexpr int KT = 64 * WK;             // k tile
  constexpr int SD = CO_T + 16, SX = KT + 16;  // LDS row strides (bf16)
  constexpr int DG = CO_T / 8, XG = KT / 8;    // 16-byte groups per pixel row
  constexpr int DC = 64 * DG / 256;       // dY groups per thread: 2 or 1
  constexpr int XC = 64 * XG / 256;       // X groups per thread: 2 or 4
  __shared__ __attribute__((aligned(16))) bf16 s_dy[64 * SD];  // [m][co]
  __shared__ __attribute__((aligned(16))) bf16 s_x[64 * SX];   // [m][k]
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  const int wco = wv % (4 / WK), wk = wv / (4 / WK);  // this wave's co strip and k half
  const int K = 9 * a.cin;
  const int co0 = blockIdx.x * CO_T, k0 = blockIdx.y * KT;
  const long long M = (long long)a.B * a.H * a.W;
  const long long m_begin = (long long)blockIdx.z * a.pix_per_split;
  const long long m_end = m_begin + a.pix_per_split < M ? m_begin + a.pix_per_split : M;

  f32x4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  // Each thread stages the same 8 k values (one tap t, channels c .. c + 7)
  // of X for XC pixels, pm = tid / XG + i (256 / XG), in every chunk; the
  // chunk's first pixel (b, y, x0) advances by 64 pixels per chunk.
  const int kx = k0 + (int)(threadIdx.x % XG) * 8;
  const int tap = kx / a.cin, cx = kx - tap * a.cin;
  const int oy = tap / 3 - 1, ox = tap % 3 - 1;
  const bool kin = kx < K;
  const int pmx = (int)threadIdx.x / XG;
  const long long hw = (long long)a.H * a.W;
  int cb = (int)(m_begin / hw);
  int cy = (int)((m_begin - cb * hw) / a.W);
  int cx0 = (int)(m_begin - cb * hw - (long long)cy * a.W);
  // Operands of one chunk in registers: the next chunk's loads are issued
  // before the current chunk's MFMAs and land in LDS after them.
  bf16x8 rd[DC], rx[XC];
  auto load = [&](long long m0) {
#pragma unroll
    for (int i = 0; i < DC; ++i) {
      const int q = threadIdx.x + 256 * i, pm = q / DG, cg = (q % DG) * 8;
      const long long m = m0 + pm;
      rd[i] = zero8();
      if (m < m_end && co0 + cg < a.cout) rd[i] = *reinterpret_cast<const bf16x8*>(a.dy + m * a.cout + co0 + cg);
    }
#pragma unroll
    for (int i = 0; i < XC; ++i) {
      const int pm = pmx + i * (256 / XG);
      rx[i] = zero8();
      if (kin && m0 + pm < m_end) {
        int x = cx0 + pm, y = cy, b = cb;
        while (x >= a.W) {
          x -= a.W;
          if (++y == a.H) {
            y = 0;
            ++b;
          }
        }
        const int yy = y + oy, xx = x + ox;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
          rx[i] = *reinterpret_cast<const bf16x8*>(conv_in(a.x, a.x2, a.c0, a.H, a.W, a.cin, b, yy, xx, cx));
      }
    }
    // the chunk's first pixel moves on by 64
    cx0 += 64;
    while (cx0 >= a.W) {
      cx0 -= a.W;
      if (++cy == a.H) {
        cy = 0;
        ++cb;
      }
    }
  };
  // Bias gradient (the blocks of the first k tile): each thread sums the dY
  // values it stages, channels (tid % DG) * 8 .. + 7, in FP32.
  const bool bias = a.bias_partial && blockIdx.y == 0;
  float bs[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (m_begin < m_end) load(m_begin);
  for (long long m0 = m_begin; m0 < m_end; m0 += 64) {
    // stage: group q -> pixel pm = q / G, channels cg .. cg + 7, cg = (q % G) * 8
#pragma unroll
    for (int i = 0; i < DC; ++i) {
      const int q = threadIdx.x + 256 * i, pm = q / DG, cg = (q % DG) * 8;
      *reinterpret_cast<bf16x8*>(&s_dy[pm * SD + cg]) = rd[i];
      if (bias) {
        const i16x8 u = __builtin_bit_cast(i16x8, rd[i]);
#pragma unroll
        for (int e = 0; e < 8; ++e) bs[e] += __uint_as_float((unsigned)(unsigned short)u[e] << 16);
      }
    }
#pragma unroll
    for (int i = 0; i < XC; ++i) {
      const int pm = pmx + i * (256 / XG);
      *reinterpret_cast<bf16x8*>(&s_x[pm * SX + (threadIdx.x % XG) * 8]) = rx[i];
    }
    __syncthreads();
    if (m0 + 64 < m_end) load(m0 + 64);
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // two 32-pixel MFMA steps
      const int row = 32 * h + 8 * g + (r16 >> 2), col = 4 * (r16 & 3);
      const bf16x8 av = tr_fragment(s_dy, row, 16 * wco + col, SD);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bf16x8 bv = tr_fragment(s_x, row, 64 * wk + 16 * j + col, SX);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* out = a.partial + (size_t)blockIdx.z * a.cout * K;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = co0 + 16 * wco + 4 * g + r, k = k0 + 64 * wk + 16 * j + r16;
      if (co < a.cout && k < K) out[(size_t)co * K + k] = acc[j][r];
    }
  if (bias) {
    // the 256 / DG threads of a channel group, added in thread order through
    // the (now idle) X tile
    float* red = reinterpret_cast<float*>(s_x);
#pragma unroll
    for (int e = 0; e < 8; ++e) red[threadIdx.x * 8 + e] = bs[e];
    __syncthreads();
    if ((int)threadIdx.x < CO_T) {
      const int grp = threadIdx.x / 8, e = threadIdx.x % 8;
      float t = 0.0f;
      for (int j = grp; j < 256; j += DG) t += red[j * 8 + e];
      if (co0 + (int)threadIdx.x < a.cout)
        a.bias_partial[(size_t)blockIdx.z * a.cout + co0 + threadIdx.x] = t;
    }
  }
}

// out[i] = sum_s partial[s][i] (fixed order)
__global__ void reduce_partials_kernel(const float* __restrict__ partial, float* __restrict__ out, long long n,
                                       int nsplit) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float p4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int k = 0;
  for (; k + 3 < nsplit; k += 4) {
#pragma unroll
    for (int j = 0; j < 4; ++j) p4[j] += partial[(size_t)(k + j) * n + i];
  }
  for (int j = 0; k < nsplit; ++k, ++j) p4[j & 3] += partial[(size_t)k * n + i];
  out[i] = (p4[0] + p4[1]) + (p4[2] + p4[3]);
}

// Column sums of an [M][C] bf16 or FP32 matrix into partial[split][C]: block
// `split` reduces rows [split rpb, split rpb + rpb).  Up to 256 columns per
// pass; the 256 threads are (256 / cw) row lanes x cw columns, so a row's
// columns are read by consecutive threads, and the row lanes are added in a
// fixed order through LDS.  blockDim.x = 256.
template <typename T>
__global__ __launch_bounds__(256) void colsum_kernel(const T* __restrict__ a, float* __restrict__ partial, long long M,
                                                     int C, long long rows_per_split) {
  __shared__ float red[256];
  const long long r0 = (long long)blockIdx.x * rows_per_split;
  const long long r1 = r0 + rows_per_split < M ? r0 + rows_per_split : M;
  for (int c0 = 0; c0 < C; c0 += 256) {
    const int cw = C - c0 < 256 ? C - c0 : 256, nrl = 256 / cw;
    const int rl = threadIdx.x / cw, c = c0 + (int)threadIdx.x % cw;
    // eight independent partial sums keep eight loads in flight
    float p8[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (rl < nrl) {
      long long r = r0 + rl;
      for (; r + 7LL * nrl < r1; r += 8LL * nrl) {
#pragma unroll
        for (int j = 0; j < 8; ++j) p8[j] += (float)a[(r + (long long)j * nrl) * C + c];
      }
      for (int j = 0; r < r1; r += nrl, ++j) p8[j & 7] += (float)a[r * C + c];
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += p8[j];
    red[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < cw) {
      float t = 0.0f;
      for (int k = 0; k < nrl; ++k) t += red[k * cw + threadIdx.x];
      partial[(size_t)blockIdx.x * C + c0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Elementwise kernels.
// ---------------------------------------------------------------------------
__global__ void pad_input_kernel(const float* __restrict__ x, bf16* __restrict__ out8, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bf16x8 v = zero8();
  v[0] = (bf16)x[i];
  *reinterpret_cast<bf16x8*>(out8 + i * 8) = v;
}

// dst[b][y][x][c] = mean of src[b][2y+i][2x+j][c]
__global__ void pool2_kernel(const bf16* __restrict__ src, bf16* __restrict__ dst, int B, int H, int W, int C) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)B * H * W * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long q = i / C;
  const int x = (int)(q % W);
  q /= W;
  const int y = (int)(q % H);
  const int b = (int)(q / H);
  const size_t row = (size_t)2 * W * C;
  const size_t o = (((size_t)b * 2 * H + 2 * y) * 2 * W + 2 * x) * C + c;
  const float s = (float)src[o] + (float)src[o + C] + (float)src[o + row] + (float)src[o + row + C];
  dst[i] = (bf16)(0.25f * s);
}

// g_src[b][2y+i][2x+j][c] += dv[b][y][x][c] / 4   (one thread per source element)
// gsrc += dv / 4 over each 2x2 block; with out (bf16) the sum, masked by
// mask > 0, goes to out instead (the last contribution to that gradient)
__global__ void pool2_back_kernel(const float* __restrict__ dv, float* __restrict__ gsrc, bf16* __restrict__ out,
                                  const bf16* __restrict__ mask, int B, int H, int W, int C) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)B * 2 * H * 2 * W * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long q = i / C;
  const int x = (int)(q % (2 * W));
  q /= 2 * W;
  const int y = (int)(q % (2 * H));
  const int b = (int)(q / (2 * H));
  const float v = gsrc[i] + 0.25f * dv[(((size_t)b * H + y / 2) * W + x / 2) * C + c];
  if (out) out[i] = (bf16)((float)mask[i] > 0.0f ? v : 0.0f);
  else gsrc[i] = v;
}

// dst[b][y][x][0:C0] = low[b][y/2][x/2][:], dst[..][C0:C0+C1] = skip[b][y][x][:]
__global__ void upcat_kernel(const bf16* __restrict__ low, const bf16* __restrict__ skip, bf16* __restrict__ dst,
                             int B, int H, int W, int C0, int C1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int C = C0 + C1;
  const long long n = (long long)B * H * W * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long long q = i / C;
  const int x = (int)(q % W);
  q /= W;
  const int y = (int)(q % H);
  const int b = (int)(q / H);
  dst[i] = c < C0 ? low[(((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C0 + c]
                  : skip[(((size_t)b * H + y) * W + x) * C1 + (c - C0)];
}

// g_low[b][y'][x'][c] = sum of the four children (c < C0); g_skip += rest
// glow = sum over each 2x2 block of dv[..., :C0]; with out (bf16), masked by
// mask > 0, into out instead
__global__ void upcat_back_low_kernel(const float* __restrict__ dv, float* __restrict__ glow, bf16* __restrict__ out,
                                      const bf16* __restrict__ mask, int B, int H, int W, int C0, int C1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Hl = H / 2, Wl = W / 2, C = C0 + C1;
  const long long n = (long long)B * Hl * Wl * C0;
  if (i >= n) return;
  const int c = (int)(i % C0);
  long long q = i / C0;
  const int x = (int)(q % Wl);
  q /= Wl;
  const int y = (int)(q % Hl);
  const int b = (int)(q / Hl);
  const size_t o = (((size_t)b * H + 2 * y) * W + 2 * x) * C + c;
  const size_t row = (size_t)W * C;
  const float v = dv[o] + dv[o + C] + dv[o + row] + dv[o + row + C];
  if (out) out[i] = (bf16)((float)mask[i] > 0.0f ? v : 0.0f);
  else glow[i] = v;
}

// gskip (+)= dv[..., C0:]
__global__ void upcat_back_skip_kernel(const float* __restrict__ dv, float* __restrict__ gskip, int B, int H, int W,
                                       int C0, int C1, int accumulate) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)B * H * W * C1;
  if (i >= n) return;
  const int c = (int)(i % C1);
  const long long p = i / C1;
  const float v = dv[p * (C0 + C1) + C0 + c];
  gskip[i] = accumulate ? gskip[i] + v : v;
}

// y[m] = x[m] + sum_c d0[m][c] w[c] + b
__global__ void head_kernel(const bf16* __restrict__ d0, const float* __restrict__ w, const float* __restrict__ b,
                            const float* __restrict__ x, float* __restrict__ y, long long M, int C) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  float s = b[0];
  if (C % 8 == 0) {
    // 16-byte loads of 8 channels
    for (int c = 0; c < C; c += 8) {
      const i16x8 u = __builtin_bit_cast(i16x8, *reinterpret_cast<const bf16x8*>(d0 + m * C + c));
#pragma unroll
      for (int e = 0; e < 8; ++e) s = __builtin_fmaf(__uint_as_float((unsigned)(unsigned short)u[e] << 16), w[c + e], s);
    }
  } else {
    for (int c = 0; c < C; ++c) s = __builtin_fmaf((float)d0[m * C + c], w[c], s);
  }
  y[m] = x[m] + s;
}

// dy = 2 (y - target) / n
// gd0 = dy w (FP32), or with out: bf16(dy w * (mask > 0)) into out.  One
// thread per (pixel, channel), so the gradient stores are coalesced; the
// channel-0 thread of a pixel also stores dy.
__global__ void head_back_kernel(const float* __restrict__ y, const float* __restrict__ target,
                                 const float* __restrict__ w, float* __restrict__ dy, float* __restrict__ gd0,
                                 bf16* __restrict__ out, const bf16* __restrict__ mask, long long M, int C,
                                 float scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * C) return;
  const long long m = i / C;
  const int c = (int)(i - m * C);
  const float d = scale * (y[m] - target[m]);
  if (c == 0) dy[m] = d;
  const float v = d * w[c];
  if (out) out[i] = (bf16)((float)mask[i] > 0.0f ? v : 0.0f);
  else gd0[i] = v;
}

// head weight gradient partials: part[s][c] = sum_m dy[m] d0[m][c], part[s][C] = sum_m dy[m]
// partial[split][c] = sum over the split's rows of dy[r] * d0[r][c] (c < C)
// or dy[r] (c = C, the bias); C + 1 <= 256 columns, row lanes as in
// colsum_kernel.  blockDim.x = 256.
__global__ __launch_bounds__(256) void head_wgrad_kernel(const float* __restrict__ dy, const bf16* __restrict__ d0,
                                                         float* __restrict__ part, long long M, int C,
                                                         long long rows_per_split) {
  __shared__ float red[256];
  const int cw = C + 1, nrl = 256 / cw;
  const int rl = threadIdx.x / cw, c = (int)threadIdx.x % cw;
  const long long r0 = (long long)blockIdx.x * rows_per_split;
  const long long r1 = r0 + rows_per_split < M ? r0 + rows_per_split : M;
  float p8[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (rl < nrl) {
    long long r = r0 + rl;
    for (; r + 7LL * nrl < r1; r += 8LL * nrl) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const long long rr = r + (long long)j * nrl;
        p8[j] += dy[rr] * (c < C ? (float)d0[rr * C + c] : 1.0f);
      }
    }
    for (int j = 0; r < r1; r += nrl, ++j) p8[j & 7] += dy[r] * (c < C ? (float)d0[r * C + c] : 1.0f);
  }
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < 8; ++j) s += p8[j];
  red[threadIdx.x] = s;
  __syncthreads();
  if ((int)threadIdx.x < cw) {
    float t = 0.0f;
    for (int k = 0; k < nrl; ++k) t += red[k * cw + threadIdx.x];
    part[(size_t)blockIdx.x * cw + threadIdx.x] = t;
  }
}

// out = bf16(g * (act > 0))
__global__ void relu_mask_kernel(const float* __restrict__ g, const bf16* __restrict__ act, bf16* __restrict__ out,
                                 long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (bf16)((float)act[i] > 0.0f ? g[i] : 0.0f);
}

// wt[c][t][o] = w[o][8 - t][c]  (input-gradient weights)
__global__ void flip_weights_kernel(const bf16* __restrict__ w, bf16* __restrict__ wt, int cin, int cout) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)cin * 9 * cout;
  if (i >= n) return;
  const int o = (int)(i % cout);
  const long long q = i / cout;
  const int t = (int)(q % 9);
  const int c = (int)(q / 9);
  wt[i] = w[((size_t)o * 9 + (8 - t)) * cin + c];
}

// Adam on an FP32 master tensor, with its bf16 copy (nullable)
__global__ void adam_kernel(float* __restrict__ w32, bf16* __restrict__ w16, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, long long n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w32[i] - lr * (mi * c1) / (__builtin_sqrtf(vi * c2) + eps);
  w32[i] = wi;
  if (w16) w16[i] = (bf16)wi;
}

__global__ void fill_kernel(float* __restrict__ p, long long n, float value) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

__global__ void to_bf16_kernel(const float* __restrict__ src, bf16* __restrict__ dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (bf16)src[i];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// 3x3 convolutions with cin % 32 == 0 (or cin == 8) and cout % 32 == 0 run
// the halo-tiled kernel; the others the K-step-staged one.
constexpr bool kHaloConv = true;

hipError_t nn_conv3x3(const void* in, const void* in2, int c0, const void* w, const float* bias, void* out,
                      const void* mask, int B, int H, int W, int cin, int cout, int relu, int out_f32, hipStream_t s) {
  ConvArgs a{static_cast<const bf16*>(in), static_cast<const bf16*>(w), bias, out, B, H, W, cin, cout, relu, out_f32,
             static_cast<const bf16*>(mask), static_cast<const bf16*>(in2), c0};
  const long long M = (long long)B * H * W;
  if (cin % 8 != 0) return hipErrorInvalidValue;
  if (kHaloConv && (cin % 32 == 0 || cin == 8) && cout % 32 == 0) {
    // halo-tiled kernel: 16-row tiles for 32 output channels, 8-row tiles
    // and 64-channel blocks otherwise
    const unsigned tiles16 = (unsigned)((W + 15) / 16), img = (unsigned)B;
    if (cout % 64 == 0) {
      const dim3 grid(tiles16 * (unsigned)((H + 7) / 8) * img, cout / 64);
      if (cin == 8) conv3x3_halo_kernel<8, 64, 2, 2, 8><<<grid, 256, 0, s>>>(a);
      else conv3x3_halo_kernel<8, 64, 2, 2, 32><<<grid, 256, 0, s>>>(a);
    } else {
      const dim3 grid(tiles16 * (unsigned)((H + 15) / 16) * img, cout / 32);
      if (cin == 8) conv3x3_halo_kernel<16, 32, 4, 1, 8><<<grid, 256, 0, s>>>(a);
      else conv3x3_halo_kernel<16, 32, 4, 1, 32><<<grid, 256, 0, s>>>(a);
    }
    return hipGetLastError();
  }
  if (cout % 128 == 0) {
    conv3x3_lds_kernel<128, 128, 2, 2><<<dim3(blocks_for(M, 128), cout / 128), 256, 0, s>>>(a);
  } else if (cout % 64 == 0) {
    conv3x3_lds_kernel<128, 64, 2, 2><<<dim3(blocks_for(M, 128), cout / 64), 256, 0, s>>>(a);
  } else if (cout % 32 == 0) {
    conv3x3_lds_kernel<256, 32, 4, 1><<<dim3(blocks_for(M, 256), cout / 32), 256, 0, s>>>(a);
  } else {
    conv3x3_kernel<4, 1><<<dim3(blocks_for(M, 128), (cout + 31) / 32), 256, 0, s>>>(a);
  }
  return hipGetLastError();
}

long long nn_wgrad_splits(long long M) {
  // about 8192 pixels per block, at least one block
  long long n = (M + 8191) / 8192;
  return n < 1 ? 1 : n;
}

// partial holds nsplit * cout * (9 cin + 1) floats: the weight partials,
// then (with db) the bias partials.
hipError_t nn_wgrad(const void* x, const void* x2, int c0, const void* dy, float* dw, float* db, float* partial, int B,
                    int H, int W, int cin, int cout, hipStream_t s) {
  const long long M = (long long)B * H * W;
  const long long ns = nn_wgrad_splits(M);
  const long long per = ((M + ns - 1) / ns + 63) / 64 * 64;
  const int K = 9 * cin;
  float* bpart = db ? partial + ns * cout * K : nullptr;
  WgradArgs a{static_cast<const bf16*>(x), static_cast<const bf16*>(dy), partial, B, H, W, cin, cout, per,
              static_cast<const bf16*>(x2), c0, bpart};
  if (cout <= 32) wgrad_kernel<32><<<dim3(1, (K + 127) / 128, (unsigned)ns), 256, 0, s>>>(a);
  else wgrad_kernel<64><<<dim3((cout + 63) / 64, (K + 63) / 64, (unsigned)ns), 256, 0, s>>>(a);
  const long long n = (long long)cout * K;
  reduce_partials_kernel<<<blocks_for(n), 256, 0, s>>>(partial, dw, n, (int)ns);
  if (db) reduce_partials_kernel<<<blocks_for(cout), 256, 0, s>>>(bpart, db, cout, (int)ns);
  return hipGetLastError();
}

hipError_t nn_colsum_bf16(const void* a, float* out, float* partial, long long M, int C, hipStream_t s) {
  const long long ns = (M + 4095) / 4096;
  colsum_kernel<bf16><<<(unsigned)ns, 256, 0, s>>>(static_cast<const bf16*>(a), partial, M, C, 4096);
  reduce_partials_kernel<<<blocks_for(C), 256, 0, s>>>(partial, out, C, (int)ns);
  return hipGetLastError();
}

hipError_t nn_pad_input(const float* x, void* out8, long long n, hipStream_t s) {
  pad_input_kernel<<<blocks_for(n), 256, 0, s>>>(x, static_cast<bf16*>(out8), n);
  return hipGetLastError();
}

hipError_t nn_pool2(const void* src, void* dst, int B, int H, int W, int C, hipStream_t s) {
  pool2_kernel<<<blocks_for((long long)B * H * W * C), 256, 0, s>>>(static_cast<const bf16*>(src),
                                                                   static_cast<bf16*>(dst), B, H, W, C);
  return hipGetLastError();
}

hipError_t nn_pool2_back(const float* dv, float* gsrc, void* out, const void* mask, int B, int H, int W, int C,
                         hipStream_t s) {
  pool2_back_kernel<<<blocks_for((long long)B * 4 * H * W * C), 256, 0, s>>>(
      dv, gsrc, static_cast<bf16*>(out), static_cast<const bf16*>(mask), B, H, W, C);
  return hipGetLastError();
}

hipError_t nn_upcat(const void* low, const void* skip, void* dst, int B, int H, int W, int C0, int C1,
                    hipStream_t s) {
  upcat_kernel<<<blocks_for((long long)B * H * W * (C0 + C1)), 256, 0, s>>>(
      static_cast<const bf16*>(low), static_cast<const bf16*>(skip), static_cast<bf16*>(dst), B, H, W, C0, C1);
  return hipGetLastError();
}

hipError_t nn_upcat_back(const float* dv, float* glow, void* glow_out, const void* glow_mask, float* gskip,
                         int skip_accumulate, int B, int H, int W, int C0, int C1, hipStream_t s) {
  upcat_back_low_kernel<<<blocks_for((long long)B * (H / 2) * (W / 2) * C0), 256, 0, s>>>(
      dv, glow, static_cast<bf16*>(glow_out), static_cast<const bf16*>(glow_mask), B, H, W, C0, C1);
  upcat_back_skip_kernel<<<blocks_for((long long)B * H * W * C1), 256, 0, s>>>(dv, gskip, B, H, W, C0, C1,
                                                                               skip_accumulate);
  return hipGetLastError();
}

hipError_t nn_head(const void* d0, const float* w, const float* b, const float* x, float* y, long long M, int C,
                   hipStream_t s) {
  head_kernel<<<blocks_for(M), 256, 0, s>>>(static_cast<const bf16*>(d0), w, b, x, y, M, C);
  return hipGetLastError();
}

hipError_t nn_head_back(const float* y, const float* target, const float* w, const void* d0, float* dy, float* gd0,
                        void* gd0_out, const void* mask, float* gw, float* partial, long long M, int C,
                        hipStream_t s) {
  head_back_kernel<<<blocks_for(M * C), 256, 0, s>>>(y, target, w, dy, gd0, static_cast<bf16*>(gd0_out),
                                                 static_cast<const bf16*>(mask), M, C, 2.0f / (float)M);
  const long long ns = (M + 4095) / 4096;
  head_wgrad_kernel<<<(unsigned)ns, 256, 0, s>>>(dy, static_cast<const bf16*>(d0), partial, M, C, 4096);
  reduce_partials_kernel<<<blocks_for(C + 1), 256, 0, s>>>(partial, gw, C + 1, (int)ns);
  return hipGetLastError();
}

hipError_t nn_relu_mask(const float* g, const void* act, void* out, long long n, hipStream_t s) {
  relu_mask_kernel<<<blocks_for(n), 256, 0, s>>>(g, static_cast<const bf16*>(act), static_cast<bf16*>(out), n);
  return hipGetLastError();
}

hipError_t nn_flip_weights(const void* w, void* wt, int cin, int cout, hipStream_t s) {
  flip_weights_kernel<<<blocks_for((long long)cin * 9 * cout), 256, 0, s>>>(static_cast<const bf16*>(w),
                                                                           static_cast<bf16*>(wt), cin, cout);
  return hipGetLastError();
}

hipError_t nn_adam(float* w32, void* w16, const float* g, float* m, float* v, long long n, float lr, float b1,
                   float b2, float eps, int step, hipStream_t s) {
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step)), c2 = 1.0f / (1.0f - powf(b2, (float)step));
  adam_kernel<<<blocks_for(n), 256, 0, s>>>(w32, static_cast<bf16*>(w16), g, m, v, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t nn_fill(float* p, long long n, float value, hipStream_t s) {
  fill_kernel<<<blocks_for(n), 256, 0, s>>>(p, n, value);
  return hipGetLastError();
}

hipError_t nn_to_bf16(const float* src, void* dst, long long n, hipStream_t s) {
  to_bf16_kernel<<<blocks_for(n), 256, 0, s>>>(src, static_cast<bf16*>(dst), n);
  return hipGetLastError();
}

}  // namespace sgdct
