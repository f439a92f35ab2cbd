// surrogate.hip -- learned surrogate for periodic 2D Gray-Scott: a residual
// CNN of five circular 3x3 convolutions (4 -> 48 -> 48 -> 48 -> 48 -> 2, GELU
// between them) that predicts the state 32 stencil steps ahead, with
// hand-written CDNA4 kernels for the forward pass, the backward pass (data and
// weight gradients), the loss and Adam.
//
//   x0  = [ (u-0.5)*2, v*4, (F-0.0375)/0.0175, (k-0.0575)/0.0075 ]   (per pixel)
//   a_l = gelu(conv_l(a_{l-1}) + b_l),  l = 1..4     (a_0 = x0)
//   z5  = conv_5(a_4) + b_5
//   [u,v](t+32) = [u,v](t) + s .* z5                (s: per-channel output scale)
//
// Numerics: weights and activations are BF16 (round to nearest even) on MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; bias, GELU, residual,
// loss and Adam in FP32 (loss reduced in FP64).  Layouts: activations NHWC
// BF16 [B][H][W][C] with C padded to a multiple of 8; the state planes FP32
// [B][2][H][W]; convolution weights BF16 [Cout][3][3][Cin] for the forward
// pass and [Cin][3][3][Cout] for the data gradient; master weights FP32.
//
// Convolution as implicit GEMM on MFMA 16x16x32: M = output channels (A =
// weights), N = pixels (B = activations), K = 9 * input channels ordered
// (dy, dx, c).  MFMA lane layout (wave64): lane l holds A[m = l%16][k =
// 8*(l/16) + j] and B[k = 8*(l/16) + j][n = l%16] for j = 0..7, and C[m =
// 4*(l/16) + i][n = l%16] for i = 0..3.  Eight consecutive k are eight
// consecutive channels of one tap, so an A or B fragment is one 16-byte load,
// and a lane's four results are four consecutive channels of one pixel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define SCHECK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      fprintf(stderr, "surrogate: HIP error '%s' at %s:%d (%s)\n", hipGetErrorString(e_), \
              __FILE__, __LINE__, #expr);                                             \
      abort();                                                                        \
    }                                                                                 \
  } while (0)

namespace sur {

typedef unsigned short bf16;
typedef short bf16x8 __attribute__((ext_vector_type(8)));        // 8 BF16 as raw bits
typedef __bf16 bf16x8_mfma __attribute__((ext_vector_type(8)));  // the MFMA operand type
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_mfma, a),
                                                 __builtin_bit_cast(bf16x8_mfma, b), c, 0, 0, 0);
}

constexpr int NL = 5;
constexpr int CIN[NL] = {4, 48, 48, 48, 48};
constexpr int COUT[NL] = {48, 48, 48, 48, 2};
constexpr int CP_IN[NL] = {8, 48, 48, 48, 48};      // first layer: 4 features x (hi, lo)
constexpr int CP_OUT[NL] = {48, 48, 48, 48, 8};     // output channels padded (layer 5: 8)
constexpr int NT_OUT[NL] = {3, 3, 3, 3, 1};         // 16-wide N tiles of the forward GEMM
static_assert(CP_IN[0] == 2 * CIN[0], "first layer: hi and lo parts of the input features");

__host__ __device__ inline bf16 f2bf(float f) {   // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16 h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// erf by Abramowitz & Stegun, Handbook of Mathematical Functions, 7.1.26:
// erf(x) = 1 - (a1 t + a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5) exp(-x^2),
// t = 1 / (1 + p |x|), |error| <= 1.5e-7 -- branch-free, a few VALU
// instructions; the activations it feeds are rounded to BF16 (8 bits).
// The polynomial is evaluated by Horner's rule with fused multiply-adds.
__device__ inline float erf_as(float x) {
  const float ax = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.0f));
  const float poly =
      fmaf(fmaf(fmaf(fmaf(1.061405429f, t, -1.453152027f), t, 1.421413741f), t, -0.284496736f), t, 0.254829592f) * t;
  return copysignf(fmaf(-poly, __expf(-ax * ax), 1.0f), x);
}
__device__ inline float gelu(float z) {
  const float hz = 0.5f * z;
  return fmaf(hz, erf_as(z * 0.70710678118654752f), hz);
}
__device__ inline float gelu_grad(float z) {
  const float cdf = fmaf(0.5f, erf_as(z * 0.70710678118654752f), 0.5f);
  const float pdf = 0.39894228040143268f * __expf(-0.5f * z * z);
  return fmaf(z, pdf, cdf);
}
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_cvt __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// two floats -> two BF16 in one 32-bit word, low half first, rounded to
// nearest even by the hardware conversion (one v_cvt_pk_bf16_f32)
__device__ inline unsigned int pack_bf16x2(float lo, float hi) {
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(f32x2{lo, hi}, bf16x2_cvt));
}
__device__ inline u32x2 pack_bf16x4(const f32x4& v) { return u32x2{pack_bf16x2(v[0], v[1]), pack_bf16x2(v[2], v[3])}; }
__device__ inline float bf16_lo(unsigned int w) { return __builtin_bit_cast(float, w << 16); }
__device__ inline float bf16_hi(unsigned int w) { return __builtin_bit_cast(float, w & 0xffff0000u); }

__device__ inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// ---------------------------------------------------------------------------
// input planes -> x0 (NHWC, 8 channels).  The four FP32 features x (u, v, F,
// k, affinely normalised) are split into hi = bf16(x) (channels 0..3) and
// lo = bf16(x - hi) (channels 4..7); the first layer applies the same weights
// to both, so it sees x with about 16 significant bits.
__device__ inline void split_bf16(float x, bf16& hi, bf16& lo) {
  hi = f2bf(x);
  lo = f2bf(x - bf2f(hi));   // x - hi is exact in FP32
}

__global__ void pack_input(const float* __restrict__ state, const float* __restrict__ fk, bf16* __restrict__ x0,
                           int B, int H, int W) {
  const size_t n = (size_t)B * H * W;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(p / ((size_t)H * W));
    const size_t hw = p % ((size_t)H * W);
    const float u = state[((size_t)b * 2 + 0) * H * W + hw];
    const float v = state[((size_t)b * 2 + 1) * H * W + hw];
    bf16* o = x0 + p * 8;
    split_bf16((u - 0.5f) * 2.0f, o[0], o[4]);
    split_bf16(v * 4.0f, o[1], o[5]);
    split_bf16((fk[2 * b] - 0.0375f) / 0.0175f, o[2], o[6]);
    split_bf16((fk[2 * b + 1] - 0.0575f) / 0.0075f, o[3], o[7]);
  }
}

// ---------------------------------------------------------------------------
// Implicit-GEMM convolution on MFMA.  One wave computes 64 consecutive pixels
// of one image row (4 M tiles) times NT output tiles of 16 channels.
//   DIR = +1: forward,       out[p][n] = sum_{d,c} wt[n][d][c] * in[p + d - 1][c]
//   DIR = -1: data gradient, out[p][n] = sum_{d,c} wt[n][d][c] * in[p - d + 1][c]
// `in` has CI channels (multiple of 8), `wt` is [NT*16 padded n][9][CI].
// Epilogues:
//   EPI_HIDDEN: z = acc (accumulated onto the bias); zq = bf16(z) -> zout;
//               a = bf16(gelu(z)) -> aout
//   EPI_OUT:    state_out = state_in + scale * acc (bias included) for n < 2
//   EPI_DGELU:  aout = bf16(acc * gelu'(zq[p][n]))   (dz of the previous layer)
enum { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_DGELU = 2, EPI_INFER = 3 };   // EPI_INFER: a_l only (no z_l kept)

// Accumulator layout: the weights are the MFMA's A operand (rows = output
// channels) and the pixels its B operand (columns), so lane (m, g) of a
// 16x16 tile holds channels 4g..4g+3 of pixel m: four consecutive channels,
// written as one 8-byte store.
//
// initial accumulator of channels c0..c0+3: their bias in the forward
// epilogues (the output layer has 2 channels), zero for the data gradient
template <int EPI>
__device__ inline f32x4 acc_init(const float* __restrict__ bias, int c0) {
  if constexpr (EPI == EPI_DGELU) {
    return f32x4{0.f, 0.f, 0.f, 0.f};
  } else if constexpr (EPI == EPI_OUT) {
    return c0 == 0 ? f32x4{bias[0], bias[1], 0.f, 0.f} : f32x4{0.f, 0.f, 0.f, 0.f};
  } else {
    return f32x4{bias[c0], bias[c0 + 1], bias[c0 + 2], bias[c0 + 3]};
  }
}

// epilogue of channels c0..c0+3 of pixel p = (b, y, x); s = the accumulator
// (bias included)
template <int EPI>
__device__ inline void epilogue4(const f32x4& s, int b, int y, int x, size_t p, int c0, int H, int W, int CO,
                                 bf16* __restrict__ zout, bf16* __restrict__ aout, const bf16* __restrict__ zq,
                                 const float* __restrict__ state_in, float* __restrict__ state_out,
                                 const float* __restrict__ scale) {
  if constexpr (EPI == EPI_HIDDEN || EPI == EPI_INFER) {
    if constexpr (EPI == EPI_HIDDEN) *(u32x2*)(zout + p * CO + c0) = pack_bf16x4(s);
    *(u32x2*)(aout + p * CO + c0) = pack_bf16x4(f32x4{gelu(s[0]), gelu(s[1]), gelu(s[2]), gelu(s[3])});
  } else if constexpr (EPI == EPI_OUT) {
    if (c0 == 0) {
      const size_t hw = (size_t)H * W;
      const size_t q = (size_t)b * 2 * hw + (size_t)y * W + x;
      state_out[q] = state_in[q] + scale[0] * s[0];
      state_out[q + hw] = state_in[q + hw] + scale[1] * s[1];
    }
  } else {
    const u32x2 zw = *(const u32x2*)(zq + p * CO + c0);
    *(u32x2*)(aout + p * CO + c0) =
        pack_bf16x4(f32x4{s[0] * gelu_grad(bf16_lo(zw[0])), s[1] * gelu_grad(bf16_hi(zw[0])),
                          s[2] * gelu_grad(bf16_lo(zw[1])), s[3] * gelu_grad(bf16_hi(zw[1]))});
  }
}

template <int CI, int NT, int DIR, int EPI>
__global__ void __launch_bounds__(256) conv_mfma(const bf16* __restrict__ in, const bf16* __restrict__ wt,
                                                 const float* __restrict__ bias, bf16* __restrict__ zout,
                                                 bf16* __restrict__ aout, const bf16* __restrict__ zq,
                                                 const float* __restrict__ state_in, float* __restrict__ state_out,
                                                 const float* __restrict__ scale, int B, int H, int W, int CO) {
  constexpr int KSTEPS = (9 * CI + 31) / 32;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int segs = W / 64;
  const int job = blockIdx.x * 4 + wave;           // (b, y, segment)
  if (job >= B * H * segs) return;
  const int b = job / (H * segs);
  const int y = (job / segs) % H;
  const int x0 = (job % segs) * 64;
  const int m = lane % 16, g = lane / 16;
  f32x4 acc[4][NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const f32x4 a0 = acc_init<EPI>(bias, nt * 16 + 4 * g);
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t][nt] = a0;
  }
  const size_t img = (size_t)b * H * W;
#pragma unroll 2
  for (int ks = 0; ks < KSTEPS; ++ks) {
    const int k = ks * 32 + g * 8;                 // first of this lane's 8 k
    const bool kvalid = k < 9 * CI;
    const int tap = kvalid ? k / CI : 0;
    const int c = k % CI;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int yy = wrapi(y + DIR * dy, H);
    bf16x8 bfr[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int n = nt * 16 + m;
      bfr[nt] = kvalid ? *(const bf16x8*)(wt + ((size_t)n * 9 + tap) * CI + c) : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int xx = wrapi(x0 + t * 16 + m + DIR * dx, W);
      const bf16x8 afr = kvalid ? *(const bf16x8*)(in + (img + (size_t)yy * W + xx) * CI + c)
                                : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        acc[t][nt] = mfma_bf16(bfr[nt], afr, acc[t][nt]);
    }
  }
  // epilogue: lane holds channels nt*16 + 4*g + i of pixel x0 + t*16 + m
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int x = x0 + t * 16 + m;
    const size_t p = img + (size_t)y * W + x;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      epilogue4<EPI>(acc[t][nt], b, y, x, p, nt * 16 + 4 * g, H, W, CO, zout, aout, zq, state_in, state_out, scale);
  }
}

// The same convolution with the input tile staged in LDS: a workgroup of 4
// waves covers 4*RPW image rows x 64 pixels; it first copies rows
// [y0-1, y0+4*RPW] x pixels [x0-1, x0+64] (periodic wrap) of all CI channels
// into LDS, then each wave computes RPW rows (4*RPW M tiles), reading its A
// fragments from LDS and each B fragment (weights, from global memory) once
// for all its M tiles.  Pixel stride in LDS: CI + PAD channels (PAD = 8: 16
// bytes of padding spread the 16 lanes of a fragment over the banks).
template <int CI, int NT, int DIR, int EPI, int RPW, int PAD>
__global__ void __launch_bounds__(256) conv_lds(const bf16* __restrict__ in, const bf16* __restrict__ wt,
                                                const float* __restrict__ bias, bf16* __restrict__ zout,
                                                bf16* __restrict__ aout, const bf16* __restrict__ zq,
                                                const float* __restrict__ state_in, float* __restrict__ state_out,
                                                const float* __restrict__ scale, int B, int H, int W, int CO) {
  constexpr int KSTEPS = (9 * CI + 31) / 32;
  constexpr int TR = 4 * RPW + 2;          // tile rows in LDS
  constexpr int TP = 66;                   // tile pixels in LDS
  constexpr int CS = CI + PAD;             // channel stride of one pixel in LDS
  __shared__ __attribute__((aligned(16))) bf16 tile[TR * TP * CS];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int segs = W / 64, bands = H / (4 * RPW);
  const int blk = blockIdx.x;              // (b, band, segment)
  const int b = blk / (bands * segs);
  const int y0 = ((blk / segs) % bands) * 4 * RPW;
  const int x0 = (blk % segs) * 64;
  const size_t img = (size_t)b * H * W;
  // stage: TR x TP pixels x CI channels, 8 channels (16 bytes) per copy; all
  // of a thread's loads are issued before its first LDS write (32-bit element
  // offsets within the image: H * W * CI < 2^31, checked at init)
  constexpr int CH8 = CI / 8;
  constexpr int NCH = TR * TP * CH8, NIT = (NCH + 255) / 256;
  const bf16* inb = in + img * CI;
  bf16x8 stg[NIT];
  int loff[NIT];
#pragma unroll
  for (int j = 0; j < NIT; ++j) {
    const int i = threadIdx.x + 256 * j;
    if (NCH % 256 == 0 || j < NIT - 1 || i < NCH) {
      const int r = i / (TP * CH8), q = i - r * (TP * CH8);
      const int px = q / CH8, c8 = q - px * CH8;
      const int yy = wrapi(y0 - 1 + r, H), xx = wrapi(x0 - 1 + px, W);
      stg[j] = *(const bf16x8*)(inb + (unsigned)((yy * W + xx) * CI + c8 * 8));
      loff[j] = (r * TP + px) * CS + c8 * 8;
    }
  }
#pragma unroll
  for (int j = 0; j < NIT; ++j)
    if (NCH % 256 == 0 || j < NIT - 1 || threadIdx.x + 256 * j < NCH) *(bf16x8*)(tile + loff[j]) = stg[j];
  __syncthreads();
  const int m = lane % 16, g = lane / 16;
  f32x4 acc[RPW][4][NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const f32x4 a0 = acc_init<EPI>(bias, nt * 16 + 4 * g);
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[r][t][nt] = a0;
  }
  // weight row n is [9][CI] = the K index k itself; the A fragment of lane
  // (m, g) at k-step ks sits at a lane base plus a per-k offset, and the
  // (row, M tile) offsets are constants
  const bf16* wrow[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) wrow[nt] = wt + (size_t)(nt * 16 + m) * 9 * CI + g * 8;
  const bf16* abase = tile + ((wave * RPW + 1) * TP + m + 1) * CS;
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ++ks) {
    const int k = ks * 32 + g * 8;
    const bool kvalid = k < 9 * CI;
    const int tap = kvalid ? k / CI : 0;
    const int c = k - tap * CI;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const bf16* ap = abase + (DIR * dy * TP + DIR * dx) * CS + c;
    bf16x8 bfr[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      bfr[nt] = kvalid ? *(const bf16x8*)(wrow[nt] + ks * 32) : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x8 afr = kvalid ? *(const bf16x8*)(ap + (r * TP + t * 16) * CS) : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[r][t][nt] = mfma_bf16(bfr[nt], afr, acc[r][t][nt]);
      }
  }
  // epilogue: lane holds channels nt*16 + 4*g + i of pixel x0 + t*16 + m
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int y = y0 + wave * RPW + r;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int x = x0 + t * 16 + m;
      const size_t p = img + (size_t)y * W + x;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        epilogue4<EPI>(acc[r][t][nt], b, y, x, p, nt * 16 + 4 * g, H, W, CO, zout, aout, zq, state_in, state_out,
                       scale);
    }
  }
}

// per instance: FP64 sums of (a - b)^2 and b^2 over its n_each floats, in a
// fixed order (block = (chunk, instance), 256 threads, then sum_pairs)
constexpr int DIFF_CHUNKS = 16;
__global__ void __launch_bounds__(256) sq_diff(const float* __restrict__ a, const float* __restrict__ b,
                                               size_t n_each, double* __restrict__ part) {
  __shared__ double sd[256], sb[256];
  const int inst = blockIdx.y, chunk = blockIdx.x;
  const size_t lo = n_each * chunk / DIFF_CHUNKS, hi = n_each * (chunk + 1) / DIFF_CHUNKS;
  const float* pa = a + (size_t)inst * n_each;
  const float* pb = b + (size_t)inst * n_each;
  double d = 0.0, r = 0.0;
  for (size_t i = lo + threadIdx.x; i < hi; i += 256) {
    const double x = pa[i], y = pb[i];
    d += (x - y) * (x - y);
    r += y * y;
  }
  sd[threadIdx.x] = d;
  sb[threadIdx.x] = r;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sd[threadIdx.x] += sd[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((size_t)inst * DIFF_CHUNKS + chunk) * 2] = sd[0];
    part[((size_t)inst * DIFF_CHUNKS + chunk) * 2 + 1] = sb[0];
  }
}

// loss and the output gradient: dz5[p][n] = 2 s_n (out - target) / count for
// n < 2 (BF16, 8 channels), per-block FP64 partial sums of (out - target)^2
__global__ void __launch_bounds__(256) loss_grad(const float* __restrict__ out, const float* __restrict__ target,
                                                 const float* __restrict__ scale, bf16* __restrict__ dz5,
                                                 double* __restrict__ part, int B, int H, int W) {
  __shared__ double red[256];
  const size_t n = (size_t)B * H * W;
  const double inv = 1.0 / (double)(n * 2);
  double acc = 0.0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(p / ((size_t)H * W));
    const size_t hw = p % ((size_t)H * W);
    bf16* o = dz5 + p * 8;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const size_t q = ((size_t)b * 2 + c) * H * W + hw;
      const float d = out[q] - target[q];
      acc += (double)d * d;
      o[c] = f2bf((float)(2.0 * inv) * d * scale[c]);
    }
#pragma unroll
    for (int c = 2; c < 8; ++c) o[c] = 0;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv;
}

// Weight gradient on MFMA: dW[n][tap][c] = sum_p dz[p][n] * a[p + d(tap)][c].
// M = output channel n (16-tiles), N = (tap, c) (16-tiles), K = pixels.  One
// wave owns one N tile for one pixel chunk (rows [r0, r0 + RC) of one image)
// and writes FP32 partial sums part[chunk][n][tap*CI + c].
template <int CI, int CO, int MT>
__global__ void __launch_bounds__(256) wgrad_mfma(const bf16* __restrict__ a, const bf16* __restrict__ dz,
                                                  float* __restrict__ part, int B, int H, int W, int RC) {
  constexpr int NCOL = 9 * CI;
  constexpr int NTILES = (NCOL + 15) / 16;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int job = blockIdx.x * 4 + wave;
  const int chunks = B * (H / RC);
  if (job >= chunks * NTILES) return;
  const int ntile = job % NTILES;
  const int chunk = job / NTILES;
  const int b = chunk / (H / RC);
  const int r0 = (chunk % (H / RC)) * RC;
  const int m = lane % 16, g = lane / 16;
  const int col = ntile * 16 + m;                   // this lane's B column (tap, c)
  const bool col_ok = col < NCOL;
  const int tap = col_ok ? col / CI : 0, c = col % CI;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  f32x4 acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const size_t img = (size_t)b * H * W;
  for (int y = r0; y < r0 + RC; ++y) {
    const int ya = wrapi(y + dy, H);
    for (int xk = 0; xk < W; xk += 32) {
      const int xs = xk + 8 * g;                    // this lane's 8 pixels: xs .. xs+7
      bf16x8 bfr;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int xa = wrapi(xs + j + dx, W);
        bfr[j] = col_ok ? (short)a[(img + (size_t)ya * W + xa) * CI + c] : (short)0;
      }
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        const int n = t * 16 + m;
        bf16x8 afr;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          afr[j] = n < CO ? (short)dz[(img + (size_t)y * W + xs + j) * (CO < 8 ? 8 : CO) + n] : (short)0;
        acc[t] = mfma_bf16(afr, bfr, acc[t]);
      }
    }
  }
  // C[m = output channel 4g+i of tile t][n = column col]
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int n = t * 16 + 4 * g + i;
      if (n < CO && col_ok) part[((size_t)chunk * CO + n) * NCOL + col] = acc[t][i];
    }
}

// bias gradient partials: part[chunk][n] = sum over the chunk's pixels of
// dz[p][n].  256 threads per chunk, one pixel per thread at a time: thread t
// sums pixels t, t + 256, ... of the chunk in that order (each pixel's CS
// channels are one run of 16-byte loads), then thread n < CO adds the 256
// threads' sums of channel n in thread order.  Fixed order, so the same bits
// on every run.
template <int CO>
__global__ void __launch_bounds__(256) bgrad(const bf16* __restrict__ dz, float* __restrict__ part, int B, int H,
                                             int W, int RC) {
  constexpr int CS = CO < 8 ? 8 : CO;   // channel stride of dz (BF16 elements): 16-byte multiples
  static_assert(CS % 8 == 0, "whole 16-byte loads per pixel");
  __shared__ float sh[256][CO + 1];
  const int chunk = blockIdx.x;
  const int b = chunk / (H / RC);
  const int r0 = (chunk % (H / RC)) * RC;
  const int P = RC * W;
  const bf16* base = dz + ((size_t)b * H + r0) * W * CS;
  float acc[CO];
#pragma unroll
  for (int n = 0; n < CO; ++n) acc[n] = 0.f;
  for (int p = threadIdx.x; p < P; p += 256) {
    const uint4* px = reinterpret_cast<const uint4*>(base + (size_t)p * CS);
#pragma unroll
    for (int v = 0; v < CS / 8; ++v) {
      const uint4 q = px[v];
      const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int n = 8 * v + j;
        if (n < CO) acc[n] += bf2f((bf16)((w4[j / 2] >> (16 * (j % 2))) & 0xffffu));
      }
    }
  }
#pragma unroll
  for (int n = 0; n < CO; ++n) sh[threadIdx.x][n] = acc[n];
  __syncthreads();
  if (threadIdx.x < CO) {
    float s = 0.f;
    for (int t = 0; t < 256; ++t) s += sh[t][threadIdx.x];
    part[(size_t)chunk * CO + threadIdx.x] = s;
  }
}

// sum partials over chunks in a fixed order: out[i] = sum_c part[c * n + i]
// as four quarter sums over consecutive chunk ranges, each in chunk order,
// added in quarter order (64 outputs x 4 quarters per 256-thread workgroup)
__global__ void __launch_bounds__(256) sum_chunks(const float* __restrict__ part, float* __restrict__ out, int chunks,
                                                  int n) {
  __shared__ float sh[4][64];
  const int t = threadIdx.x & 63, k = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + t;
  const int c0 = (int)((long long)chunks * k / 4), c1 = (int)((long long)chunks * (k + 1) / 4);
  float s = 0.f;
  if (i < n)
    for (int c = c0; c < c1; ++c) s += part[(size_t)c * n + i];
  sh[k][t] = s;
  __syncthreads();
  if (k == 0 && i < n) out[i] = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
}
inline int sum_chunks_blocks(size_t n) { return (int)((n + 63) / 64); }

// x += sigma * N(0,1) with a counter-based hash (Box-Muller over two uniforms):
// input-noise augmentation for training
__global__ void add_noise(float* __restrict__ x, size_t n, float sigma, unsigned long long seed) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ULL * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const float u1 = ((float)(z >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = ((float)((z >> 16) & 0xffffffULL)) * (1.0f / 16777216.0f);
    x[i] += sigma * sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
  }
}

// Adam on the FP32 master parameters
__global__ void adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                     float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    w[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
  }
}

// master [co][tap][ci] FP32 -> forward BF16 [co_pad][tap][ci_pad] and data-gradient
// BF16 [ci_pad_n][tap][co_pad_k] (zero padded)
// with tie != 0 (first layer, cif == 2 * ci), padded input channel c >= ci
// takes the weight of channel c - ci (the lo parts of the split input)
__global__ void cast_weights(const float* __restrict__ w, bf16* __restrict__ wf, bf16* __restrict__ wb, int ci,
                             int co, int cif, int cof, int cib, int cob, int tie) {
  const int n = cof * 9 * cif;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int c = i % cif;
    const int tap = (i / cif) % 9, o = i / (cif * 9);
    if (tie && c >= ci) c -= ci;
    wf[i] = (o < co && c < ci) ? f2bf(w[((size_t)o * 9 + tap) * ci + c]) : (bf16)0;
  }
  const int nb = cib * 9 * cob;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
    const int o = i % cob, tap = (i / cob) % 9, c = i / (cob * 9);
    wb[i] = (o < co && c < ci) ? f2bf(w[((size_t)o * 9 + tap) * ci + c]) : (bf16)0;
  }
}

// first-layer weight gradient: g[o][tap][c] = s[o][tap][c] + s[o][tap][c + ci]
// (s over the 2 * ci split input channels, whose two halves share weights)
__global__ void fold_tied(const float* __restrict__ s, float* __restrict__ g, int co, int ci) {
  const int n = co * 9 * ci;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = i % ci, ot = i / ci;
    g[i] = s[(size_t)ot * 2 * ci + c] + s[(size_t)ot * 2 * ci + c + ci];
  }
}

// ---------------------------------------------------------------------------
struct Net {
  int H = 0, W = 0, maxB = 0;
  int RC = 16;                      // rows per weight-gradient chunk
  // parameters: per layer W [co][9][ci] then b [co], FP32, contiguous
  int off_w[NL], off_b[NL], nparam = 0;
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  bf16 *wf[NL], *wb[NL];
  float* scale = nullptr;           // output scale s (2)
  bf16 *x0 = nullptr, *z[4], *a[4], *dz5 = nullptr, *dza = nullptr, *dzb = nullptr;
  float *part = nullptr, *out = nullptr;
  double* lpart = nullptr;
  double* dpart = nullptr;   // sq_diff partial sums: maxB x DIFF_CHUNKS x 2
  size_t part_floats = 0;
  hipStream_t stream = nullptr;
  int adam_t = 0;
  bool ready = false;
};

Net net;

int grid_for(size_t n, int threads) {
  size_t b = (n + threads - 1) / threads;
  return (int)(b < 65536 ? b : 65536);
}

void cast_all() {
  for (int l = 0; l < NL; ++l) {
    const int cob = (l < NL - 1) ? 48 : 8;    // data-gradient K channels (layer 5: dz has 8)
    const int cib = 48;                       // data-gradient N (input channels, padded to 3 tiles)
    hipLaunchKernelGGL(cast_weights, dim3(64), dim3(256), 0, net.stream, net.w + net.off_w[l], net.wf[l], net.wb[l],
                       CIN[l], COUT[l], CP_IN[l], NT_OUT[l] * 16, cib, cob, l == 0 ? 1 : 0);
  }
  SCHECK(hipGetLastError());
}

// convolution kernel: 0 = fragments from global memory (conv_mfma), 1 = LDS
// tile of 4 rows, 2 = LDS tile of 8 rows (conv_lds; padded pixels; default)
static int conv_kind = 2;

template <int CI, int NT, int DIR, int EPI>
void launch_conv(const bf16* in, const bf16* wt, const float* bias, bf16* z, bf16* a, const bf16* zq,
                 const float* s_in, float* s_out, const float* scale, int B, int H, int W, int CO) {
  if (conv_kind == 2 && H % 8 == 0) {
    hipLaunchKernelGGL((conv_lds<CI, NT, DIR, EPI, 2, 8>), dim3(B * (H / 8) * (W / 64)), dim3(256), 0, net.stream,
                       in, wt, bias, z, a, zq, s_in, s_out, scale, B, H, W, CO);
  } else if (conv_kind == 1 && H % 4 == 0) {
    hipLaunchKernelGGL((conv_lds<CI, NT, DIR, EPI, 1, 8>), dim3(B * (H / 4) * (W / 64)), dim3(256), 0, net.stream,
                       in, wt, bias, z, a, zq, s_in, s_out, scale, B, H, W, CO);
  } else {
    hipLaunchKernelGGL((conv_mfma<CI, NT, DIR, EPI>), dim3((B * H * (W / 64) + 3) / 4), dim3(256), 0, net.stream,
                       in, wt, bias, z, a, zq, s_in, s_out, scale, B, H, W, CO);
  }
}

// train: keep the pre-activations z_l for the backward pass
void forward(const float* state, const float* fk, float* out_state, int B, bool train = true) {
  const int H = net.H, W = net.W;
  const size_t px = (size_t)B * H * W;
  hipLaunchKernelGGL(pack_input, dim3(grid_for(px, 256)), dim3(256), 0, net.stream, state, fk, net.x0, B, H, W);
  const bf16* in = net.x0;
  if (train) {
    launch_conv<8, 3, 1, EPI_HIDDEN>(in, net.wf[0], net.w + net.off_b[0], net.z[0], net.a[0], nullptr, nullptr,
                                     nullptr, nullptr, B, H, W, 48);
    for (int l = 1; l < 4; ++l)
      launch_conv<48, 3, 1, EPI_HIDDEN>(net.a[l - 1], net.wf[l], net.w + net.off_b[l], net.z[l], net.a[l], nullptr,
                                        nullptr, nullptr, nullptr, B, H, W, 48);
  } else {
    launch_conv<8, 3, 1, EPI_INFER>(in, net.wf[0], net.w + net.off_b[0], nullptr, net.a[0], nullptr, nullptr,
                                    nullptr, nullptr, B, H, W, 48);
    for (int l = 1; l < 4; ++l)
      launch_conv<48, 3, 1, EPI_INFER>(net.a[l - 1], net.wf[l], net.w + net.off_b[l], nullptr, net.a[l], nullptr,
                                       nullptr, nullptr, nullptr, B, H, W, 48);
  }
  launch_conv<48, 1, 1, EPI_OUT>(net.a[3], net.wf[4], net.w + net.off_b[4], nullptr, nullptr, nullptr, state,
                                 out_state, net.scale, B, H, W, 8);
  SCHECK(hipGetLastError());
}

// gradients of the loss (out vs target) into net.g; returns the loss
double backward(const float* target, int B) {
  const int H = net.H, W = net.W;
  const size_t px = (size_t)B * H * W;
  const int lblocks = 1024;
  hipLaunchKernelGGL(loss_grad, dim3(lblocks), dim3(256), 0, net.stream, net.out, target, net.scale, net.dz5,
                     net.lpart, B, H, W);
  const int chunks = B * (H / net.RC);
  const int blocks = (int)((B * H * (W / 64) + 3) / 4);
  // layer 5: weights from a4 and dz5; data gradient -> dz4
  const bf16* dz = net.dz5;
  bf16* dzn = net.dza;
  for (int l = NL - 1; l >= 0; --l) {
    const bf16* ain = l == 0 ? net.x0 : net.a[l - 1];
    const int ncol = 9 * CIN[l];
    const int ci = CP_IN[l];
    // weight-gradient partials
    const int ntiles = (9 * ci + 15) / 16;
    const int wblocks = (chunks * ntiles + 3) / 4;
    if (l == NL - 1)
      hipLaunchKernelGGL((wgrad_mfma<48, 2, 1>), dim3(wblocks), dim3(256), 0, net.stream, ain, dz, net.part, B, H, W,
                         net.RC);
    else if (l == 0)
      hipLaunchKernelGGL((wgrad_mfma<8, 48, 3>), dim3(wblocks), dim3(256), 0, net.stream, ain, dz, net.part, B, H, W,
                         net.RC);
    else
      hipLaunchKernelGGL((wgrad_mfma<48, 48, 3>), dim3(wblocks), dim3(256), 0, net.stream, ain, dz, net.part, B, H,
                         W, net.RC);
    // the first layer's partials hold columns for the 8 split input channels:
    // sum the chunks, then fold the hi and lo halves (tied weights)
    const int ncol_pad = 9 * ci;
    float* sum = net.part + (size_t)chunks * COUT[l] * ncol_pad;
    hipLaunchKernelGGL(sum_chunks, dim3(sum_chunks_blocks((size_t)COUT[l] * ncol_pad)), dim3(256), 0, net.stream,
                       net.part, sum, chunks, COUT[l] * ncol_pad);
    if (ci == CIN[l]) {
      SCHECK(hipMemcpyAsync(net.g + net.off_w[l], sum, sizeof(float) * COUT[l] * ncol, hipMemcpyDeviceToDevice,
                            net.stream));
    } else {
      hipLaunchKernelGGL(fold_tied, dim3(grid_for((size_t)COUT[l] * ncol, 256)), dim3(256), 0, net.stream, sum,
                         net.g + net.off_w[l], COUT[l], CIN[l]);
    }
    // bias gradient
    if (COUT[l] == 2)
      hipLaunchKernelGGL((bgrad<2>), dim3(chunks), dim3(256), 0, net.stream, dz, net.part, B, H, W, net.RC);
    else
      hipLaunchKernelGGL((bgrad<48>), dim3(chunks), dim3(256), 0, net.stream, dz, net.part, B, H, W, net.RC);
    hipLaunchKernelGGL(sum_chunks, dim3(sum_chunks_blocks(COUT[l])), dim3(256), 0, net.stream, net.part,
                       net.g + net.off_b[l], chunks,
                       COUT[l]);
    // data gradient into the previous layer (times gelu')
    if (l > 0) {
      if (l == NL - 1)
        launch_conv<8, 3, -1, EPI_DGELU>(dz, net.wb[l], nullptr, nullptr, dzn, net.z[l - 1], nullptr, nullptr, nullptr,
                                         B, H, W, 48);
      else
        launch_conv<48, 3, -1, EPI_DGELU>(dz, net.wb[l], nullptr, nullptr, dzn, net.z[l - 1], nullptr, nullptr,
                                          nullptr, B, H, W, 48);
      dz = dzn;
      dzn = (dzn == net.dza) ? net.dzb : net.dza;
    }
    (void)px;
  }
  SCHECK(hipGetLastError());
  std::vector<double> lp(lblocks);
  SCHECK(hipMemcpyAsync(lp.data(), net.lpart, sizeof(double) * lblocks, hipMemcpyDeviceToHost, net.stream));
  SCHECK(hipStreamSynchronize(net.stream));
  double loss = 0.0;
  for (double v : lp) loss += v;
  return loss;
}

}  // namespace sur

using namespace sur;

extern "C" {

int Sur_Num_Params(void) {
  int n = 0;
  for (int l = 0; l < NL; ++l) n += COUT[l] * 9 * CIN[l] + COUT[l];
  return n;
}

// H, W: image size (W a multiple of 64, H a multiple of 16); maxB: largest batch
int Sur_Init(int H, int W, int maxB) {
  // the convolutions address one image with 32-bit element offsets
  if (net.ready || W % 64 != 0 || H % 16 != 0 || maxB < 1 || (long long)H * W * 48 >= (1LL << 31)) return -1;
  net.H = H;
  net.W = W;
  net.maxB = maxB;
  SCHECK(hipStreamCreateWithFlags(&net.stream, hipStreamNonBlocking));
  int off = 0;
  for (int l = 0; l < NL; ++l) {
    net.off_w[l] = off;
    off += COUT[l] * 9 * CIN[l];
    net.off_b[l] = off;
    off += COUT[l];
  }
  net.nparam = off;
  for (float** p : {&net.w, &net.g, &net.m1, &net.m2}) {
    SCHECK(hipMalloc(p, sizeof(float) * off));
    SCHECK(hipMemsetAsync(*p, 0, sizeof(float) * off, net.stream));
  }
  for (int l = 0; l < NL; ++l) {
    SCHECK(hipMalloc(&net.wf[l], sizeof(bf16) * NT_OUT[l] * 16 * 9 * CP_IN[l]));
    SCHECK(hipMalloc(&net.wb[l], sizeof(bf16) * 48 * 9 * 48));
  }
  SCHECK(hipMalloc(&net.scale, sizeof(float) * 2));
  const size_t px = (size_t)maxB * H * W;
  SCHECK(hipMalloc(&net.x0, sizeof(bf16) * px * 8));
  for (int l = 0; l < 4; ++l) {
    SCHECK(hipMalloc(&net.z[l], sizeof(bf16) * px * 48));
    SCHECK(hipMalloc(&net.a[l], sizeof(bf16) * px * 48));
  }
  SCHECK(hipMalloc(&net.dz5, sizeof(bf16) * px * 8));
  SCHECK(hipMalloc(&net.dza, sizeof(bf16) * px * 48));
  SCHECK(hipMalloc(&net.dzb, sizeof(bf16) * px * 48));
  SCHECK(hipMalloc(&net.out, sizeof(float) * px * 2));
  const int chunks = maxB * (H / net.RC);
  net.part_floats = (size_t)chunks * 48 * 9 * 48 + (size_t)48 * 9 * 48;
  SCHECK(hipMalloc(&net.part, sizeof(float) * net.part_floats));
  SCHECK(hipMalloc(&net.lpart, sizeof(double) * 1024));
  SCHECK(hipMalloc(&net.dpart, sizeof(double) * maxB * DIFF_CHUNKS * 2));
  const float one[2] = {1.f, 1.f};
  SCHECK(hipMemcpyAsync(net.scale, one, sizeof(one), hipMemcpyHostToDevice, net.stream));
  SCHECK(hipStreamSynchronize(net.stream));
  net.ready = true;
  return 0;
}

int Sur_Set_Params(const float* host) {
  if (!net.ready) return -1;
  SCHECK(hipMemcpyAsync(net.w, host, sizeof(float) * net.nparam, hipMemcpyHostToDevice, net.stream));
  cast_all();
  SCHECK(hipMemsetAsync(net.m1, 0, sizeof(float) * net.nparam, net.stream));
  SCHECK(hipMemsetAsync(net.m2, 0, sizeof(float) * net.nparam, net.stream));
  net.adam_t = 0;
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

int Sur_Get_Params(float* host) {
  if (!net.ready) return -1;
  SCHECK(hipMemcpyAsync(host, net.w, sizeof(float) * net.nparam, hipMemcpyDeviceToHost, net.stream));
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

int Sur_Get_Grads(float* host) {
  if (!net.ready) return -1;
  SCHECK(hipMemcpyAsync(host, net.g, sizeof(float) * net.nparam, hipMemcpyDeviceToHost, net.stream));
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

int Sur_Set_Scale(const float* s2) {
  if (!net.ready) return -1;
  SCHECK(hipMemcpyAsync(net.scale, s2, sizeof(float) * 2, hipMemcpyHostToDevice, net.stream));
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

// convolution kernel choice (0: fragments from global memory, 1: LDS tile of
// 4 rows (default), 2: LDS tile of 8 rows; all compute the same sums in the
// same order)
int Sur_Set_Conv(int kind) {
  if (kind < 0 || kind > 2) return -1;
  conv_kind = kind;
  return 0;
}

// relative L2 distance of a from b per instance (B instances of n_each
// floats each, device pointers): rel[i] = ||a_i - b_i|| / ||b_i|| (FP64 sums)
int Sur_Rel_L2(const float* a, const float* b, int B, size_t n_each, double* rel) {
  if (!net.ready || B < 1 || B > net.maxB) return -1;
  hipLaunchKernelGGL(sq_diff, dim3(DIFF_CHUNKS, B), dim3(256), 0, net.stream, a, b, n_each, net.dpart);
  SCHECK(hipGetLastError());
  std::vector<double> part((size_t)B * DIFF_CHUNKS * 2);
  SCHECK(hipMemcpyAsync(part.data(), net.dpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, net.stream));
  SCHECK(hipStreamSynchronize(net.stream));
  for (int i = 0; i < B; ++i) {
    double d = 0.0, r = 0.0;
    for (int c = 0; c < DIFF_CHUNKS; ++c) {
      d += part[((size_t)i * DIFF_CHUNKS + c) * 2];
      r += part[((size_t)i * DIFF_CHUNKS + c) * 2 + 1];
    }
    rel[i] = sqrt(d) / (r > 0.0 ? sqrt(r) : 1e-300);
  }
  return 0;
}

// state, fk, out: device pointers. state/out [B][2][H][W] FP32, fk [B][2] FP32
int Sur_Forward(const float* state, const float* fk, float* out, int B) {
  if (!net.ready || B < 1 || B > net.maxB) return -1;
  forward(state, fk, out, B, false);
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

// forward + loss + gradients (no update); loss written to *loss
int Sur_Gradients(const float* state, const float* fk, const float* target, int B, double* loss) {
  if (!net.ready || B < 1 || B > net.maxB) return -1;
  forward(state, fk, net.out, B);
  *loss = backward(target, B);
  return 0;
}

// one Adam step (gradients from Sur_Gradients or computed here)
int Sur_Train_Step(const float* state, const float* fk, const float* target, int B, float lr, double* loss) {
  if (!net.ready || B < 1 || B > net.maxB) return -1;
  forward(state, fk, net.out, B);
  *loss = backward(target, B);
  net.adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)net.adam_t), c2 = 1.f - powf(b2, (float)net.adam_t);
  hipLaunchKernelGGL(adam, dim3(grid_for(net.nparam, 256)), dim3(256), 0, net.stream, net.w, net.g, net.m1, net.m2,
                     net.nparam, lr, b1, b2, 1e-8f, c1, c2);
  cast_all();
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

// device memory helpers, so a caller needs no other GPU runtime binding
void* Sur_Alloc(size_t bytes) {
  void* p = nullptr;
  SCHECK(hipMalloc(&p, bytes));
  return p;
}
int Sur_Free(void* p) {
  SCHECK(hipFree(p));
  return 0;
}
int Sur_Copy(void* dst, const void* src, size_t bytes, int kind) {   // 0 H2D, 1 D2H, 2 D2D
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost
                                                                        : hipMemcpyDeviceToDevice);
  SCHECK(hipMemcpy(dst, src, bytes, k));
  return 0;
}

// dst[i * n_each .. (i+1) * n_each) = *ptrs[i] for i < nptr (device pointers
// listed in host memory); one launch-free copy per pointer on the library stream
int Sur_Gather(const unsigned long long* host_ptrs, float* dst, int nptr, size_t n_each) {
  for (int i = 0; i < nptr; ++i)
    SCHECK(hipMemcpyAsync(dst + (size_t)i * n_each, (const void*)host_ptrs[i], n_each * sizeof(float),
                          hipMemcpyDeviceToDevice, net.stream));
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

// *ptrs[i] = src[i * n_each .. (i+1) * n_each) for i < nptr: the inverse of
// Sur_Gather, asynchronous copies on the library stream and one wait at the end
int Sur_Scatter(const unsigned long long* host_ptrs, const float* src, int nptr, size_t n_each) {
  if (nptr < 0) return -1;
  for (int i = 0; i < nptr; ++i)
    SCHECK(hipMemcpyAsync((void*)host_ptrs[i], src + (size_t)i * n_each, n_each * sizeof(float),
                          hipMemcpyDeviceToDevice, net.stream));
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

int Sur_Add_Noise(float* x, size_t n, float sigma, unsigned long long seed) {
  hipLaunchKernelGGL(add_noise, dim3(grid_for(n, 256)), dim3(256), 0, net.stream, x, n, sigma, seed);
  SCHECK(hipGetLastError());
  SCHECK(hipStreamSynchronize(net.stream));
  return 0;
}

// time `reps` forward passes of batch B on device buffers; returns ms (events)
float Sur_Time_Forward(const float* state, const float* fk, float* out, int B, int reps) {
  if (!net.ready || B < 1 || B > net.maxB) return -1.f;
  hipEvent_t e0, e1;
  SCHECK(hipEventCreate(&e0));
  SCHECK(hipEventCreate(&e1));
  SCHECK(hipEventRecord(e0, net.stream));
  for (int r = 0; r < reps; ++r) forward(state, fk, out, B, false);
  SCHECK(hipEventRecord(e1, net.stream));
  SCHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  SCHECK(hipEventElapsedTime(&ms, e0, e1));
  SCHECK(hipEventDestroy(e0));
  SCHECK(hipEventDestroy(e1));
  return ms;
}

}  // extern "C"
