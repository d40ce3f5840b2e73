// Neural-network kernels of the Stokes warm-start U-Net (include/ps_nn.h).
//
// BF16 storage, FP32 accumulation. Activations are NHWC (channels fastest), so a 3x3
// convolution is the product of im2col(x) (rows = pixels, columns = (ky, kx, c)) with the
// weights stored [9*Cin][Cout]; the GEMM gathers the im2col operand from x as it loads its
// tiles (ps_nn_gemm_im2col), or reads a stored matrix (ps_nn_gemm). It runs on the gfx950 matrix cores:
// v_mfma_f32_16x16x32_bf16, lane l holding A[l%16][(l/16)*8 .. +7], B[(l/16)*8 .. +7][l%16] and
// D[(l/16)*4 + r][l%16], r = 0..3. A workgroup of four waves computes a 64x64 tile, each wave a
// 32x32 quarter as 2x2 MFMA blocks, over K steps of 32 staged through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "ps_amd.h"
#include "ps_internal.h"
#include "ps_nn.h"

namespace {

typedef unsigned short bf16_t;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(short)))) short s16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {  // round to nearest even, NaN -> 0x7fc0
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;

// C[m, n] = sum_k A(m, k) B(k, n), B(k, n) = B[k*sbk + n*sbn]; A(m, k) = A[m*sam + k*sak] (AM 0),
// or A is the 3x3 im2col of the NHWC input A (cH x cW x cC per image) without the column matrix:
// AM 1 A(m, k) = col[m][k] (m a pixel, k = tap*cC + c), AM 2 A(m, k) = col[k][m].
// blockIdx.z splits K into chunks of kchunk (a multiple of TK).
template <int EPI, int AM>
__global__ void __launch_bounds__(256) nn_gemm_k(int M, int N, int K, const bf16_t* __restrict__ A, long long sam, long long sak,
                                                 const bf16_t* __restrict__ B, long long sbk, long long sbn, void* __restrict__ C,
                                                 long long ldc, const float* __restrict__ bias, int kchunk, int cH, int cW, int cC) {
  __shared__ __attribute__((aligned(16))) bf16_t As[TM * LDK];
  __shared__ __attribute__((aligned(16))) bf16_t Bs[TN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const long long m0 = (long long)blockIdx.x * TM, n0 = (long long)blockIdx.y * TN;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // Each thread moves one 8-element piece of each 64x32 tile per K step: 8 consecutive k of one
  // row when k has unit stride, else 8 consecutive rows of one k (unit stride along m / n).
  // Whole, aligned pieces travel as one 16-byte load; edge or strided pieces element by element.
  // The next step's pieces are loaded while the current tile is multiplied.
  const bool a_kfast = AM == 0 ? sak == 1 : AM == 1, b_kfast = sbk == 1;
  const int ar = a_kfast ? tid / 4 : (tid % 8) * 8, ac = a_kfast ? (tid % 4) * 8 : tid / 8;
  const int br = b_kfast ? tid / 4 : (tid % 8) * 8, bc = b_kfast ? (tid % 4) * 8 : tid / 8;
  // A piece starts at an element index that is a multiple of 8 when the stride across pieces is
  // (k0, ac, ar, m0 are all multiples of 8), so the 16-byte load is aligned
  const bool a_vec = AM == 0 ? (a_kfast ? sam : sak) % 8 == 0 && ((uintptr_t)A & 15) == 0 && (a_kfast || sam == 1)
                             : cC % 8 == 0 && ((uintptr_t)A & 15) == 0;
  // implicit im2col: the input element that column j of pixel p reads (0 outside the image)
  // Every call has p < M or p < K and j < K or j < M (all int), so the 32-bit unsigned quotients
  // are exact; only the address needs 64 bits.
  auto col_at = [&](long long p, int j, int& tap, int& c, long long& src) -> bool {
    const unsigned pu = (unsigned)p, ju = (unsigned)j;
    const unsigned tu = ju / (unsigned)cC;
    tap = (int)tu;
    c = (int)(ju - tu * (unsigned)cC);
    const unsigned t2 = pu / (unsigned)cW;
    const int xx = (int)(pu - t2 * (unsigned)cW);
    const unsigned b = t2 / (unsigned)cH;
    const int yy = (int)(t2 - b * (unsigned)cH);
    const int sy = yy + tap / 3 - 1, sx = xx + tap % 3 - 1;
    src = (((long long)b * cH + sy) * cW + sx) * cC + c;
    return sy >= 0 && sy < cH && sx >= 0 && sx < cW;
  };
  const bool b_vec = (b_kfast ? sbn : sbk) % 8 == 0 && ((uintptr_t)B & 15) == 0 && (b_kfast || sbn == 1);
  s16x8 pa, pb;
  auto fetch = [&](int k0) {
    {  // A piece
      const long long m = m0 + ar;
      const int k = k0 + ac;
      const bool full = a_kfast ? (m < M && k + 7 < ke) : (m + 7 < M && k < ke);
      if (AM == 0) {
        if (a_vec && full)
          pa = *(const s16x8*)(A + m * sam + (long long)k * sak);
        else
          for (int t = 0; t < 8; ++t) {
            const long long mm = a_kfast ? m : m + t;
            const int kk = a_kfast ? k + t : k;
            pa[t] = (mm < M && kk < ke) ? (short)A[mm * sam + (long long)kk * sak] : (short)0;
          }
      } else {
        // the piece is columns j..j+7 of pixel p (AM 1: p = m, j = k; AM 2: p = k, j = m); with cC a
        // multiple of 8 they are 8 channels of one tap, one 16-byte load (j is a multiple of 8)
        const long long p = AM == 1 ? m : (long long)k;
        const int j = AM == 1 ? k : (int)m;
        int tap, c;
        long long src;
        if (a_vec && full) {
          if (col_at(p, j, tap, c, src)) pa = *(const s16x8*)(A + src);
          else pa = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
        } else {
          for (int t = 0; t < 8; ++t) {
            const bool in = AM == 1 ? (m < M && k + t < ke) : (m + t < M && k < ke);
            pa[t] = in && col_at(p, j + t, tap, c, src) ? (short)A[src] : (short)0;
          }
        }
      }
    }
    {  // B piece
      const long long n = n0 + br;
      const int k = k0 + bc;
      const bool full = b_kfast ? (n < N && k + 7 < ke) : (n + 7 < N && k < ke);
      if (b_vec && full)
        pb = *(const s16x8*)(B + (long long)k * sbk + n * sbn);
      else
        for (int t = 0; t < 8; ++t) {
          const long long nn = b_kfast ? n : n + t;
          const int kk = b_kfast ? k + t : k;
          pb[t] = (nn < N && kk < ke) ? (short)B[(long long)kk * sbk + nn * sbn] : (short)0;
        }
    }
  };
  fetch(kb);
  for (int k0 = kb; k0 < ke; k0 += TK) {
    // tiles K-fastest in LDS
    if (a_kfast) {
      *(s16x8*)&As[ar * LDK + ac] = pa;
    } else {
      for (int t = 0; t < 8; ++t) As[(ar + t) * LDK + ac] = (bf16_t)pa[t];
    }
    if (b_kfast) {
      *(s16x8*)&Bs[br * LDK + bc] = pb;
    } else {
      for (int t = 0; t < 8; ++t) Bs[(br + t) * LDK + bc] = (bf16_t)pb[t];
    }
    __syncthreads();
    if (k0 + TK < ke) fetch(k0 + TK);
    const int i = lane & 15, kq = (lane >> 4) * 8;
    s16x8 a[2], b[2];
    for (int q = 0; q < 2; ++q) {  // rows are LDK * 2 = 80 bytes apart, kq * 2 a multiple of 16
      a[q] = *(const s16x8*)&As[(wm * 32 + q * 16 + i) * LDK + kq];
      b[q] = *(const s16x8*)&Bs[(wn * 32 + q * 16 + i) * LDK + kq];
    }
    for (int bi = 0; bi < 2; ++bi)
      for (int bj = 0; bj < 2; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[bi]), __builtin_bit_cast(bf16x8, b[bj]),
                                                              acc[bi][bj], 0, 0, 0);
    __syncthreads();
  }
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int r = 0; r < 4; ++r) {
        const long long m = m0 + wm * 32 + bi * 16 + (lane >> 4) * 4 + r;
        const long long n = n0 + wn * 32 + bj * 16 + (lane & 15);
        if (m >= M || n >= N) continue;
        float v = acc[bi][bj][r];
        if (EPI == PS_NN_EPI_BF16) {
          ((bf16_t*)C)[m * ldc + n] = f2bf(v);
        } else if (EPI == PS_NN_EPI_BF16_BIAS_RELU) {
          v += bias[n];
          ((bf16_t*)C)[m * ldc + n] = f2bf(v > 0.f ? v : 0.f);
        } else if (EPI == PS_NN_EPI_BF16_BIAS) {
          ((bf16_t*)C)[m * ldc + n] = f2bf(v + bias[n]);
        } else if (EPI == PS_NN_EPI_F32_ADD) {
          atomicAdd(&((float*)C)[m * ldc + n], v);
        } else {
          ((float*)C)[m * ldc + n] = v;
        }
      }
}

unsigned grid_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 1 << 20); }

#define GS_LOOP(e, n) for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < (n); e += (long long)gridDim.x * 256)

__global__ void __launch_bounds__(256) nn_im2col_k(const bf16_t* __restrict__ x, int B, int H, int W, int C, bf16_t* __restrict__ col) {
  const long long n = (long long)B * H * W * 9 * C;
  GS_LOOP(e, n) {
    const int c = (int)(e % C);
    const long long t = e / C;
    const int kk = (int)(t % 9);
    const long long p = t / 9;
    const int xx = (int)(p % W);
    const long long t2 = p / W;
    const int yy = (int)(t2 % H), b = (int)(t2 / H);
    const int sy = yy + kk / 3 - 1, sx = xx + kk % 3 - 1;
    col[e] = (sy >= 0 && sy < H && sx >= 0 && sx < W) ? x[(((long long)b * H + sy) * W + sx) * C + c] : (bf16_t)0;
  }
}

// dx[b, y, x, c] = sum over the 3x3 taps of the columns that read (b, y, x, c)
__global__ void __launch_bounds__(256) nn_col2im_k(const bf16_t* __restrict__ dcol, int B, int H, int W, int C, bf16_t* __restrict__ dx) {
  const long long n = (long long)B * H * W * C;
  GS_LOOP(e, n) {
    const int c = (int)(e % C);
    const long long p = e / C;
    const int xx = (int)(p % W);
    const long long t2 = p / W;
    const int yy = (int)(t2 % H), b = (int)(t2 / H);
    float s = 0.f;
    for (int kk = 0; kk < 9; ++kk) {
      const int oy = yy - (kk / 3 - 1), ox = xx - (kk % 3 - 1);  // output pixel whose tap kk reads (yy, xx)
      if (oy < 0 || oy >= H || ox < 0 || ox >= W) continue;
      s += bf2f(dcol[((((long long)b * H + oy) * W + ox) * 9 + kk) * C + c]);
    }
    dx[e] = f2bf(s);
  }
}

__global__ void __launch_bounds__(256) nn_pool2_k(const bf16_t* __restrict__ x, int B, int H, int W, int C, bf16_t* __restrict__ y) {
  const int Ho = H / 2, Wo = W / 2;
  const long long n = (long long)B * Ho * Wo * C;
  GS_LOOP(e, n) {
    const int c = (int)(e % C);
    const long long p = e / C;
    const int X = (int)(p % Wo);
    const long long t = p / Wo;
    const int Y = (int)(t % Ho), b = (int)(t / Ho);
    const long long base = (((long long)b * H + 2 * Y) * W + 2 * X) * C + c;
    const float s = (bf2f(x[base]) + bf2f(x[base + C])) + (bf2f(x[base + (long long)W * C]) + bf2f(x[base + (long long)W * C + C]));
    y[e] = f2bf(0.25f * s);
  }
}

__global__ void __launch_bounds__(256) nn_pool2_bwd_k(const bf16_t* __restrict__ dy, int B, int H, int W, int C, bf16_t* __restrict__ dx) {
  const long long n = (long long)B * H * W * C;
  GS_LOOP(e, n) {
    const int c = (int)(e % C);
    const long long p = e / C;
    const int xx = (int)(p % W);
    const long long t = p / W;
    const int yy = (int)(t % H), b = (int)(t / H);
    dx[e] = f2bf(0.25f * bf2f(dy[(((long long)b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c]));
  }
}

__global__ void __launch_bounds__(256) nn_up2_k(const bf16_t* __restrict__ x, int B, int H, int W, int C, bf16_t* __restrict__ y) {
  const int Ho = 2 * H, Wo = 2 * W;
  const long long n = (long long)B * Ho * Wo * C;
  GS_LOOP(e, n) {
    const int c = (int)(e % C);
    const long long p = e / C;
    const int X = (int)(p % Wo);
    const long long t = p / Wo;
    const int Y = (int)(t % Ho), b = (int)(t / Ho);
    y[e] = x[(((long long)b * H + Y / 2) * W + X / 2) * C + c];
  }
}

__global__ void __launch_bounds__(256) nn_up2_bwd_k(const bf16_t* __restrict__ dy, int B, int H, int W, int C, bf16_t* __restrict__ dx) {
  const long long n = (long long)B * H * W * C;
  const int Wo = 2 * W;
  GS_LOOP(e, n) {
    const int c = (int)(e % C);
    const long long p = e / C;
    const int xx = (int)(p % W);
    const long long t = p / W;
    const int yy = (int)(t % H), b = (int)(t / H);
    const long long base = (((long long)b * 2 * H + 2 * yy) * Wo + 2 * xx) * C + c;
    const float s = (bf2f(dy[base]) + bf2f(dy[base + C])) + (bf2f(dy[base + (long long)Wo * C]) + bf2f(dy[base + (long long)Wo * C + C]));
    dx[e] = f2bf(s);
  }
}

__global__ void __launch_bounds__(256) nn_concat_k(const bf16_t* __restrict__ a, int Ca, const bf16_t* __restrict__ b, int Cb, long long P,
                                                   bf16_t* __restrict__ y) {
  const int C = Ca + Cb;
  GS_LOOP(e, P * C) {
    const int c = (int)(e % C);
    const long long p = e / C;
    y[e] = c < Ca ? a[p * Ca + c] : b[p * Cb + (c - Ca)];
  }
}

__global__ void __launch_bounds__(256) nn_split_k(const bf16_t* __restrict__ dy, int Ca, int Cb, long long P, bf16_t* __restrict__ da,
                                                  bf16_t* __restrict__ db) {
  const int C = Ca + Cb;
  GS_LOOP(e, P * C) {
    const int c = (int)(e % C);
    const long long p = e / C;
    if (c < Ca) da[p * Ca + c] = dy[e];
    else db[p * Cb + (c - Ca)] = dy[e];
  }
}

__global__ void __launch_bounds__(256) nn_relu_bwd_k(bf16_t* __restrict__ dy, const bf16_t* __restrict__ y, long long n) {
  GS_LOOP(e, n) {
    if (!(bf2f(y[e]) > 0.f)) dy[e] = 0;
  }
}

// db[n] += sum over rows of dy[m, n] (FP32); rows split over blockIdx.x in chunks of 256
__global__ void __launch_bounds__(256) nn_bias_grad_k(const bf16_t* __restrict__ dy, long long M, int N, float* __restrict__ db) {
  const long long r0 = (long long)blockIdx.x * 256, r1 = std::min<long long>(M, r0 + 256);
  for (int n = threadIdx.x; n < N; n += 256) {
    float s = 0.f;
    for (long long m = r0; m < r1; ++m) s += bf2f(dy[m * N + n]);
    atomicAdd(&db[n], s);
  }
}

// loss += scale * mask * (pred - target)^2 ; dpred = 2 * scale * mask * (pred - target)
__global__ void __launch_bounds__(256) nn_mse_k(const bf16_t* __restrict__ pred, const float* __restrict__ target, const float* __restrict__ mask,
                                                long long n, float scale, bf16_t* __restrict__ dpred, float* __restrict__ loss) {
  __shared__ float part[256];
  float s = 0.f;
  GS_LOOP(e, n) {
    const float d = bf2f(pred[e]) - target[e];
    const float mk = mask ? mask[e] : 1.f;
    s += scale * mk * d * d;
    dpred[e] = f2bf(2.f * scale * mk * d);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

__global__ void __launch_bounds__(256) nn_adam_k(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                                                 long long n, float lr, float b1, float b2, float eps, float c1, float c2,
                                                 bf16_t* __restrict__ w16) {
  GS_LOOP(e, n) {
    const float ge = g[e];
    const float me = b1 * m[e] + (1.f - b1) * ge;
    const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    const float we = w[e] - lr * (me * c1) / (sqrtf(ve * c2) + eps);
    w[e] = we;
    if (w16) w16[e] = f2bf(we);
  }
}

__global__ void __launch_bounds__(256) nn_f32_to_bf16_k(const float* __restrict__ x, long long n, bf16_t* __restrict__ y) {
  GS_LOOP(e, n) y[e] = f2bf(x[e]);
}
__global__ void __launch_bounds__(256) nn_bf16_to_f32_k(const bf16_t* __restrict__ x, long long n, float* __restrict__ y) {
  GS_LOOP(e, n) y[e] = bf2f(x[e]);
}
__global__ void __launch_bounds__(256) nn_fill_f32_k(float* __restrict__ x, long long n, float v) {
  GS_LOOP(e, n) x[e] = v;
}
// the weights of a 3x3 convolution's data gradient: wf[(t*Cout + co)*Cin + ci] = w[((8-t)*Cin + ci)*Cout + co]
// (taps reversed, each tap's Cin x Cout block transposed), so that dx = im2col(dy) . wf
__global__ void __launch_bounds__(256) nn_conv_flip_k(const bf16_t* __restrict__ w, int Cin, int Cout, bf16_t* __restrict__ wf) {
  GS_LOOP(e, 9LL * Cin * Cout) {
    const int ci = (int)(e % Cin);
    const long long r = e / Cin;
    const int co = (int)(r % Cout), t = (int)(r / Cout);
    wf[e] = w[((long long)(8 - t) * Cin + ci) * Cout + co];
  }
}
__global__ void __launch_bounds__(256) nn_add_bf16_k(bf16_t* __restrict__ y, const bf16_t* __restrict__ x, long long n) {
  GS_LOOP(e, n) y[e] = f2bf(bf2f(y[e]) + bf2f(x[e]));
}

int launched(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ps_internal_fail(PS_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

}  // namespace

#define NN_BEGIN                                   \
  if (int r_ = ps_internal_require_init()) return r_; \
  hipStream_t st = ps_internal_stream(stream);

namespace {

template <int AM>
int gemm_launch(int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B, int64_t sbk, int64_t sbn, void* C, int64_t ldc,
                const float* bias, int epi, int splitk, hipStream_t st, int cH, int cW, int cC) {
  if (M < 0 || N < 0 || K < 0 || splitk < 1) return ps_internal_fail(PS_EARGUMENT, "ps_nn_gemm: negative size or splitk < 1");
  if (M == 0 || N == 0) return 0;
  if ((epi == PS_NN_EPI_BF16_BIAS_RELU || epi == PS_NN_EPI_BF16_BIAS) && !bias) return ps_internal_fail(PS_EARGUMENT, "ps_nn_gemm: bias missing");
  if (splitk > 1 && epi != PS_NN_EPI_F32_ADD) return ps_internal_fail(PS_EARGUMENT, "ps_nn_gemm: split-K needs the accumulating FP32 epilogue");
  int kchunk = (K + splitk - 1) / splitk;
  kchunk = ((kchunk + TK - 1) / TK) * TK;
  if (kchunk < TK) kchunk = TK;
  const int nz = (K + kchunk - 1) / kchunk;
  dim3 grid((unsigned)((M + TM - 1) / TM), (unsigned)((N + TN - 1) / TN), (unsigned)std::max(1, nz));
  if (grid.y > 65535 || grid.z > 65535) return ps_internal_fail(PS_EARGUMENT, "ps_nn_gemm: N or split too large");
  const bf16_t* a = (const bf16_t*)A;
  const bf16_t* b = (const bf16_t*)B;
#define NN_GEMM_CASE(E)                                                                                                        \
  case E:                                                                                                                      \
    hipLaunchKernelGGL((nn_gemm_k<E, AM>), grid, dim3(256), 0, st, M, N, K, a, sam, sak, b, sbk, sbn, C, ldc, bias, kchunk, cH, cW, cC); \
    break;
  switch (epi) {
    NN_GEMM_CASE(PS_NN_EPI_BF16)
    NN_GEMM_CASE(PS_NN_EPI_BF16_BIAS_RELU)
    NN_GEMM_CASE(PS_NN_EPI_BF16_BIAS)
    NN_GEMM_CASE(PS_NN_EPI_F32_ADD)
    NN_GEMM_CASE(PS_NN_EPI_F32)
    default: return ps_internal_fail(PS_EARGUMENT, "ps_nn_gemm: unknown epilogue");
  }
#undef NN_GEMM_CASE
  return launched("ps_nn_gemm");
}

}  // namespace

extern "C" int ps_nn_gemm(int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B, int64_t sbk, int64_t sbn, void* C,
                          int64_t ldc, const float* bias, int epi, int splitk, void* stream) {
  NN_BEGIN
  return gemm_launch<0>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, epi, splitk, st, 1, 1, 1);
}

extern "C" int ps_nn_gemm_im2col(int transpose, int M, int N, int K, const void* x, int H, int W, int Cin, const void* B, int64_t sbk, int64_t sbn,
                                 void* C, int64_t ldc, const float* bias, int epi, int splitk, void* stream) {
  NN_BEGIN
  if (H < 1 || W < 1 || Cin < 1) return ps_internal_fail(PS_EARGUMENT, "ps_nn_gemm_im2col: H, W and C must be positive");
  const long long npix = transpose ? (long long)K : (long long)M, ncol = transpose ? (long long)M : (long long)K;
  if (ncol != 9LL * Cin || npix % ((long long)H * W) != 0)
    return ps_internal_fail(PS_EARGUMENT, "ps_nn_gemm_im2col: the im2col operand is (B*H*W) x (9*C): sizes do not match");
  return transpose ? gemm_launch<2>(M, N, K, x, 0, 0, B, sbk, sbn, C, ldc, bias, epi, splitk, st, H, W, Cin)
                   : gemm_launch<1>(M, N, K, x, 0, 0, B, sbk, sbn, C, ldc, bias, epi, splitk, st, H, W, Cin);
}

// Element counts are products of the sizes: a pair of negative sizes would make a positive count
// for indices the kernels never expect, so any negative size is refused before the launch.
inline bool nn_negative(std::initializer_list<long long> v) {
  for (long long x : v)
    if (x < 0) return true;
  return false;
}
#define NN_SIZES(what, ...) \
  if (nn_negative({__VA_ARGS__})) return ps_internal_fail(PS_EARGUMENT, what ": negative size");

#define NN_ELEMENTWISE(name, n, ...)                                            \
  NN_BEGIN                                                                      \
  if ((n) <= 0) return 0;                                                       \
  hipLaunchKernelGGL(name, dim3(grid_for(n)), dim3(256), 0, st, __VA_ARGS__); \
  return launched(#name);

extern "C" int ps_nn_im2col(const void* x, int B, int H, int W, int C, void* col, void* stream) {
  NN_SIZES("ps_nn_im2col", B, H, W, C)
  NN_ELEMENTWISE(nn_im2col_k, (long long)B * H * W * 9 * C, (const bf16_t*)x, B, H, W, C, (bf16_t*)col)
}
extern "C" int ps_nn_conv_flip(const void* w, int Cin, int Cout, void* wf, void* stream) {
  NN_SIZES("ps_nn_conv_flip", Cin, Cout)
  NN_ELEMENTWISE(nn_conv_flip_k, 9LL * Cin * Cout, (const bf16_t*)w, Cin, Cout, (bf16_t*)wf)
}
extern "C" int ps_nn_col2im(const void* dcol, int B, int H, int W, int C, void* dx, void* stream) {
  NN_SIZES("ps_nn_col2im", B, H, W, C)
  NN_ELEMENTWISE(nn_col2im_k, (long long)B * H * W * C, (const bf16_t*)dcol, B, H, W, C, (bf16_t*)dx)
}
extern "C" int ps_nn_pool2(const void* x, int B, int H, int W, int C, void* y, void* stream) {
  NN_SIZES("ps_nn_pool2", B, H, W, C)
  if ((H | W) & 1) return ps_internal_fail(PS_EARGUMENT, "ps_nn_pool2: H and W must be even");
  NN_ELEMENTWISE(nn_pool2_k, (long long)B * (H / 2) * (W / 2) * C, (const bf16_t*)x, B, H, W, C, (bf16_t*)y)
}
extern "C" int ps_nn_pool2_bwd(const void* dy, int B, int H, int W, int C, void* dx, void* stream) {
  NN_SIZES("ps_nn_pool2_bwd", B, H, W, C)
  if ((H | W) & 1) return ps_internal_fail(PS_EARGUMENT, "ps_nn_pool2_bwd: H and W must be even");
  NN_ELEMENTWISE(nn_pool2_bwd_k, (long long)B * H * W * C, (const bf16_t*)dy, B, H, W, C, (bf16_t*)dx)
}
extern "C" int ps_nn_up2(const void* x, int B, int H, int W, int C, void* y, void* stream) {
  NN_SIZES("ps_nn_up2", B, H, W, C)
  NN_ELEMENTWISE(nn_up2_k, (long long)B * 4 * H * W * C, (const bf16_t*)x, B, H, W, C, (bf16_t*)y)
}
extern "C" int ps_nn_up2_bwd(const void* dy, int B, int H, int W, int C, void* dx, void* stream) {
  NN_SIZES("ps_nn_up2_bwd", B, H, W, C)
  NN_ELEMENTWISE(nn_up2_bwd_k, (long long)B * H * W * C, (const bf16_t*)dy, B, H, W, C, (bf16_t*)dx)
}
extern "C" int ps_nn_concat(const void* a, int Ca, const void* b, int Cb, int64_t P, void* y, void* stream) {
  NN_SIZES("ps_nn_concat", Ca, Cb, P)
  NN_ELEMENTWISE(nn_concat_k, (long long)P * (Ca + Cb), (const bf16_t*)a, Ca, (const bf16_t*)b, Cb, (long long)P, (bf16_t*)y)
}
extern "C" int ps_nn_split(const void* dy, int Ca, int Cb, int64_t P, void* da, void* db, void* stream) {
  NN_SIZES("ps_nn_split", Ca, Cb, P)
  NN_ELEMENTWISE(nn_split_k, (long long)P * (Ca + Cb), (const bf16_t*)dy, Ca, Cb, (long long)P, (bf16_t*)da, (bf16_t*)db)
}
extern "C" int ps_nn_relu_bwd(void* dy, const void* y, int64_t n, void* stream) {
  NN_ELEMENTWISE(nn_relu_bwd_k, (long long)n, (bf16_t*)dy, (const bf16_t*)y, (long long)n)
}
extern "C" int ps_nn_bias_grad(const void* dy, int64_t M, int N, float* db, void* stream) {
  NN_BEGIN
  if (M <= 0 || N <= 0) return 0;
  hipLaunchKernelGGL(nn_bias_grad_k, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, (const bf16_t*)dy, (long long)M, N, db);
  return launched("ps_nn_bias_grad");
}
extern "C" int ps_nn_mse(const void* pred, const float* target, const float* mask, int64_t n, float scale, void* dpred, float* loss,
                         void* stream) {
  NN_ELEMENTWISE(nn_mse_k, (long long)n, (const bf16_t*)pred, target, mask, (long long)n, scale, (bf16_t*)dpred, loss)
}
extern "C" int ps_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                          void* w16, void* stream) {
  NN_ELEMENTWISE(nn_adam_k, (long long)n, w, g, m, v, (long long)n, lr, b1, b2, eps, c1, c2, (bf16_t*)w16)
}
extern "C" int ps_nn_f32_to_bf16(const float* x, int64_t n, void* y, void* stream) {
  NN_ELEMENTWISE(nn_f32_to_bf16_k, (long long)n, x, (long long)n, (bf16_t*)y)
}
extern "C" int ps_nn_bf16_to_f32(const void* x, int64_t n, float* y, void* stream) {
  NN_ELEMENTWISE(nn_bf16_to_f32_k, (long long)n, (const bf16_t*)x, (long long)n, y)
}
extern "C" int ps_nn_fill_f32(float* x, int64_t n, float v, void* stream) { NN_ELEMENTWISE(nn_fill_f32_k, (long long)n, x, (long long)n, v) }
extern "C" int ps_nn_add_bf16(void* y, const void* x, int64_t n, void* stream) {
  NN_ELEMENTWISE(nn_add_bf16_k, (long long)n, (bf16_t*)y, (const bf16_t*)x, (long long)n)
}

extern "C" int ps_nn_alloc(int64_t bytes, void** out) {
  if (int r = ps_internal_require_init()) return r;
  if (bytes < 0) return ps_internal_fail(PS_EARGUMENT, "ps_nn_alloc: negative size");
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, (size_t)std::max<int64_t>(bytes, 256));
  if (e != hipSuccess) return ps_internal_fail(e == hipErrorOutOfMemory ? PS_EOOM : PS_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemset(p, 0, (size_t)std::max<int64_t>(bytes, 256));
  if (e != hipSuccess) return ps_internal_fail(PS_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
  *out = p;
  return 0;
}
extern "C" int ps_nn_free(void* p) {
  if (!p) return 0;
  hipError_t e = hipFree(p);
  if (e != hipSuccess) return ps_internal_fail(PS_EHIP, std::string("hipFree: ") + hipGetErrorString(e));
  return 0;
}
extern "C" int ps_nn_h2d(void* dst, const void* src, int64_t bytes, void* stream) {
  NN_BEGIN
  hipError_t e = hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return ps_internal_fail(PS_EHIP, std::string("ps_nn_h2d: ") + hipGetErrorString(e));
  return 0;
}
extern "C" int ps_nn_d2h(void* dst, const void* src, int64_t bytes, void* stream) {
  NN_BEGIN
  hipError_t e = hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return ps_internal_fail(PS_EHIP, std::string("ps_nn_d2h: ") + hipGetErrorString(e));
  return 0;
}
extern "C" int ps_nn_d2d(void* dst, const void* src, int64_t bytes, void* stream) {
  NN_BEGIN
  hipError_t e = hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return ps_internal_fail(PS_EHIP, std::string("ps_nn_d2d: ") + hipGetErrorString(e));
  return 0;
}
