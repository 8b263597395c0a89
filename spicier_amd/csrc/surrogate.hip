// surrogate.hip — the learned Newton warm start: an MLP
//   n_in -> H -> H -> n_out, SiLU hidden activations,
// trained and evaluated entirely by this file's kernels.
//
// Matmuls: one tiled GEMM on v_mfma_f32_16x16x16_bf16 (64x64 block tile,
// four waves each owning a 32x32 quarter as 2x2 MFMA tiles, K staged through
// LDS 16 at a time).  Operands are FP32 in HBM and rounded to BF16
// (round-to-nearest-even) as they are staged, accumulation is FP32; strides
// describe either operand transposed, which covers the forward pass and both
// backward products.  The epilogue fuses bias add and SiLU.
// Everything else (standardisation, loss, SiLU backward, bias gradients,
// Adam) is elementwise.  FP32 master weights; Adam with bias correction and a
// cosine learning-rate decay.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

#include "handle.h"

using namespace spx;

struct spx_surrogate {
  spx_handle* h = nullptr;
  int n_in = 0, H = 0, n_out = 0;
  int P = 0;
  std::vector<int> slots;
  int* d_slots = nullptr;
  float* d_w = nullptr;     // FP32 master weights W1,b1,W2,b2,W3,b3
  float* d_grad = nullptr;
  float* d_m = nullptr;     // Adam moments
  float* d_v = nullptr;
  float* d_norm = nullptr;  // in_mean[n_in], in_std[n_in], out_mean[n_out], out_std[n_out]
  float* d_wout = nullptr;  // per-output loss weights (training only)
  int64_t nw = 0;
  int64_t adam_t = 0;
  // activation buffers, sized for `cap` rows
  int64_t cap = 0;
  float *X = nullptr, *T = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr,
        *Y = nullptr, *dY = nullptr, *dH = nullptr, *dZ = nullptr, *d_loss = nullptr;
  int* d_perm = nullptr;
  int64_t perm_cap = 0;
  // BF16 copies of W1, W2, W3 for the fused inference kernel, transposed to
  // [out][in] with the input dimension padded to a multiple of 16
  unsigned short* d_wt = nullptr;
  int64_t wt_cap = 0;
  // the FP32 weights change with every version bump (creation, set_weights,
  // each training step); d_wt holds the BF16 copy of version wt_version
  uint64_t w_version = 1, wt_version = 0;

  int64_t off_W1() const { return 0; }
  int64_t off_b1() const { return (int64_t)n_in * H; }
  int64_t off_W2() const { return off_b1() + H; }
  int64_t off_b2() const { return off_W2() + (int64_t)H * H; }
  int64_t off_W3() const { return off_b2() + H; }
  int64_t off_b3() const { return off_W3() + (int64_t)H * n_out; }
};

namespace spx {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float v) {
  uint32_t b = __float_as_uint(v);
  b += 0x7fffu + ((b >> 16) & 1u);  // round to nearest even
  return (unsigned short)(b >> 16);
}

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }

constexpr int TM = 64, TN = 64, TK = 16;

// C[m][n] = sum_k A(m,k) B(k,n) (+ beta*C) (+ bias[n]); optional H = silu(C).
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], C row-major with ldc.
__global__ __launch_bounds__(256) void k_gemm_bf16(const float* __restrict__ A, int64_t sam,
                                                   int64_t sak, const float* __restrict__ B,
                                                   int64_t sbk, int64_t sbn, float* C,
                                                   int64_t ldc, int M, int N, int K, float beta,
                                                   const float* __restrict__ bias,
                                                   float* __restrict__ Hout) {
  __shared__ unsigned short As[TM][TK + 2];
  __shared__ unsigned short Bs[TK][TN + 2];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += TK) {
    for (int q = 0; q < 4; ++q) {
      const int e = tid + q * 256;
      const int r = e / TK, c = e % TK;  // A tile 64 x 16
      const int m = m0 + r, k = k0 + c;
      As[r][c] = to_bf16((m < M && k < K) ? A[(int64_t)m * sam + (int64_t)k * sak] : 0.f);
      const int rb = e / TN, cb = e % TN;  // B tile 16 x 64
      const int kb = k0 + rb, nb = n0 + cb;
      Bs[rb][cb] = to_bf16((kb < K && nb < N) ? B[(int64_t)kb * sbk + (int64_t)nb * sbn] : 0.f);
    }
    __syncthreads();
    const int r16 = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      bf16x4 a;
      for (int t = 0; t < 4; ++t) a[t] = (short)As[wm + i * 16 + r16][kq + t];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bf16x4 b;
        for (int t = 0; t < 4; ++t) b[t] = (short)Bs[kq + t][wn + j * 16 + r16];
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const int c16 = lane & 15, rq = (lane >> 4) * 4;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + wm + i * 16 + rq + t, n = n0 + wn + j * 16 + c16;
        if (m < M && n < N) {
          float v = acc[i][j][t];
          float* cp = C + (int64_t)m * ldc + n;
          if (beta != 0.f) v += beta * *cp;
          if (bias) v += bias[n];
          *cp = v;
          if (Hout) Hout[(int64_t)m * ldc + n] = v * sigmoidf(v);
        }
      }
}

// X[r][i] = (params[row][slot_i] - mean_i) / std_i,  row = perm ? perm[base+r] : base+r
__global__ void k_gather_inputs(const double* __restrict__ params, int P,
                                const int* __restrict__ slots, int n_in,
                                const float* __restrict__ norm, const int* __restrict__ perm,
                                int64_t base, int64_t rows, float* __restrict__ X) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * n_in) return;
  int64_t r = t / n_in;
  int i = (int)(t - r * n_in);
  int64_t row = perm ? perm[base + r] : base + r;
  X[t] = (float)((params[row * P + slots[i]] - norm[i]) / norm[n_in + i]);
}

// T[r][j] = (x[row][j] - out_mean_j) / out_std_j
__global__ void k_gather_targets(const double* __restrict__ x, int n_x, int n_out, int n_in,
                                 const float* __restrict__ norm, const int* __restrict__ perm,
                                 int64_t base, int64_t rows, float* __restrict__ T) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * n_out) return;
  int64_t r = t / n_out;
  int j = (int)(t - r * n_out);
  int64_t row = perm ? perm[base + r] : base + r;
  const float* om = norm + 2 * n_in;
  const float* os = om + n_out;
  T[t] = (float)((x[row * n_x + j] - om[j]) / os[j]);
}

// dY = 2 w (Y - T) / count; loss += sum w (Y - T)^2 / count, w per output
// column (1 when wout is null)
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int64_t count,
                      int n_out, const float* __restrict__ wout, float* __restrict__ dY,
                      float* __restrict__ loss) {
  __shared__ float red[256];
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (t < count) {
    const float w = wout ? wout[t % n_out] : 1.f;
    float d = Y[t] - T[t];
    dY[t] = 2.f * w * d / (float)count;
    e = w * d * d / (float)count;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// dZ = dH * silu'(Z),  silu'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void k_silu_backward(const float* __restrict__ dH, const float* __restrict__ Z,
                                int64_t count, float* __restrict__ dZ) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  float z = Z[t], s = sigmoidf(z);
  dZ[t] = dH[t] * s * (1.f + z * (1.f - s));
}

// db[n] = sum over rows of dZ[row][n]
__global__ void k_colsum(const float* __restrict__ dZ, int64_t rows, int N,
                         float* __restrict__ db) {
  __shared__ float red[256];
  const int n = blockIdx.x;
  float s = 0.f;
  for (int64_t r = threadIdx.x; r < rows; r += 256) s += dZ[r * N + n];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[n] = red[0];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  float gt = g[t];
  float mt = b1 * m[t] + (1.f - b1) * gt;
  float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  w[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// x0[r][j] = Y[r][j] * out_std_j + out_mean_j for j < n_out, 0 beyond
__global__ void k_write_prediction(const float* __restrict__ Y, int n_out, int n_x, int n_in,
                                   const float* __restrict__ norm, int64_t base, int64_t rows,
                                   double* __restrict__ x0) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * n_x) return;
  int64_t r = t / n_x;
  int j = (int)(t - r * n_x);
  const float* om = norm + 2 * n_in;
  const float* os = om + n_out;
  x0[(base + r) * n_x + j] = j < n_out ? (double)Y[r * n_out + j] * os[j] + om[j] : 0.0;
}

// ------------------------------------------------------- fused inference
// The forward pass of one block of 64 rows in one workgroup: inputs gathered
// and standardised straight from the sweep parameters, the three layers on
// v_mfma_f32_16x16x16_bf16 with the hidden activations kept in LDS as BF16,
// and the prediction written de-standardised as FP64.  Every rounding step is
// the layered path's (inputs rounded to BF16 as the GEMM stages them,
// accumulation over K in steps of 16 in order, bias then SiLU in FP32, BF16
// rounding of the activation as the next GEMM stages it), so both paths give
// the same x0.  Weights are read as BF16 [out][in] straight from L2: each
// workgroup streams them once.
constexpr int kInfRows = 64;

// Wt[n][k] = bf16(W[k][n]) for n < N, k < K; zero up to np rows of kp
// columns (the padding), so no separate clear is needed
__global__ void k_weights_to_bf16t(const float* __restrict__ W, int K, int N, int kp, int np,
                                   unsigned short* __restrict__ Wt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)np * kp) return;
  const int n = (int)(t / kp), k = (int)(t - (int64_t)n * kp);
  Wt[t] = (k < K && n < N) ? to_bf16(W[(int64_t)k * N + n]) : (unsigned short)0;
}

// out[r][c] for c in this wave's columns: act(in[r][:] . Wt[c][:] + bias[c])
// over K (multiple of 16); `in` is LDS BF16 with row stride ldi.  Each wave
// takes ncol/4 columns; cols per wave must be a multiple of 16 (NT tiles).
template <int NT, int K, class Epi>
__device__ __forceinline__ void infer_layer(const unsigned short* in, int ldi,
                                            const unsigned short* __restrict__ Wt, int col0,
                                            int lane, Epi epi) {
  floatx4 acc[4][NT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int r16 = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll 4
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 b[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const unsigned short* wp = Wt + (int64_t)(col0 + j * 16 + r16) * K + k0 + kq;
      const uint2 w = *reinterpret_cast<const uint2*>(wp);
      b[j][0] = (short)(w.x & 0xffffu);
      b[j][1] = (short)(w.x >> 16);
      b[j][2] = (short)(w.y & 0xffffu);
      b[j][3] = (short)(w.y >> 16);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint2 av = *reinterpret_cast<const uint2*>(in + (i * 16 + r16) * ldi + k0 + kq);
      bf16x4 a;
      a[0] = (short)(av.x & 0xffffu);
      a[1] = (short)(av.x >> 16);
      a[2] = (short)(av.y & 0xffffu);
      a[3] = (short)(av.y >> 16);
#pragma unroll
      for (int j = 0; j < NT; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[j], acc[i][j], 0, 0, 0);
    }
  }
  const int c16 = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) epi(i * 16 + rq + t, col0 + j * 16 + c16, acc[i][j][t]);
}

// hidden width = 64 * HT (each wave owns 16 * HT columns), inputs padded to KP
template <int HT, int KP>
__global__ __launch_bounds__(256, 2) void k_mlp_infer(
    const double* __restrict__ params, int P, const int* __restrict__ slots, int n_in,
    const float* __restrict__ norm, const unsigned short* __restrict__ W1t,
    const float* __restrict__ b1, const unsigned short* __restrict__ W2t,
    const float* __restrict__ b2, const unsigned short* __restrict__ W3t,
    const float* __restrict__ b3, int n_out, int n_x, int64_t B, double* __restrict__ x0) {
  constexpr int H = 64 * HT;
  constexpr int LDH = H + 8;  // BF16 row stride (pads LDS banks)
  constexpr int kp = KP, ldx = KP + 8;
  extern __shared__ unsigned short lds[];
  unsigned short* Xs = lds;                  // [64][ldx]
  unsigned short* H1 = Xs + kInfRows * ldx;  // [64][LDH]
  unsigned short* H2 = H1 + kInfRows * LDH;  // [64][LDH]
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const float* om = norm + 2 * n_in;
  const float* os = om + n_out;
  for (int64_t r0 = (int64_t)blockIdx.x * kInfRows; r0 < B; r0 += (int64_t)gridDim.x * kInfRows) {
    // inputs (k_gather_inputs' arithmetic), padded columns zero
    for (int e = tid; e < kInfRows * kp; e += 256) {
      const int r = e / kp, i = e - r * kp;
      const int64_t row = r0 + r;
      float v = 0.f;
      if (i < n_in && row < B) v = (float)((params[row * P + slots[i]] - norm[i]) / norm[n_in + i]);
      Xs[r * ldx + i] = to_bf16(v);
    }
    __syncthreads();
    infer_layer<HT, KP>(Xs, ldx, W1t, wave * 16 * HT, lane, [&](int r, int c, float a) {
      const float v = a + b1[c];
      H1[r * LDH + c] = to_bf16(v * sigmoidf(v));
    });
    __syncthreads();
    infer_layer<HT, H>(H1, LDH, W2t, wave * 16 * HT, lane, [&](int r, int c, float a) {
      const float v = a + b2[c];
      H2[r * LDH + c] = to_bf16(v * sigmoidf(v));
    });
    __syncthreads();
    // output layer: n_out <= 64, one 16-column tile per wave
    infer_layer<1, H>(H2, LDH, W3t, wave * 16, lane, [&](int r, int c, float a) {
      const int64_t row = r0 + r;
      if (row < B && c < n_out) x0[row * n_x + c] = (double)(a + b3[c]) * os[c] + om[c];
    });
    for (int e = tid; e < kInfRows * (n_x - n_out); e += 256) {
      const int r = e / (n_x - n_out), j = n_out + e % (n_x - n_out);
      if (r0 + r < B) x0[(r0 + r) * n_x + j] = 0.0;
    }
    __syncthreads();
  }
}

// Column means / standard deviations over the masked rows (FP64 sums).
__global__ void k_colstats(const double* __restrict__ src, int64_t ld, const int* __restrict__ cols,
                           int64_t B, const uint8_t* __restrict__ mask, float* __restrict__ mean,
                           float* __restrict__ stdv) {
  __shared__ double s1[256], s2[256], sc[256];
  const int c = blockIdx.x;
  const int col = cols ? cols[c] : c;
  double a = 0, b = 0, k = 0;
  for (int64_t r = threadIdx.x; r < B; r += 256) {
    if (mask && !mask[r]) continue;
    double v = src[r * ld + col];
    a += v;
    b += v * v;
    k += 1;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  sc[threadIdx.x] = k;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
      sc[threadIdx.x] += sc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double cnt = sc[0] > 0 ? sc[0] : 1.0;
    double mu = s1[0] / cnt;
    double var = s2[0] / cnt - mu * mu;
    double sd = var > 0 ? sqrt(var) : 0.0;
    mean[c] = (float)mu;
    stdv[c] = (float)(sd > 1e-6 * (fabs(mu) + 1e-30) && sd > 1e-12 ? sd : 1.0);
  }
}

inline unsigned blocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

void gemm(hipStream_t st, const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk,
          int64_t sbn, float* C, int64_t ldc, int M, int N, int K, float beta, const float* bias,
          float* Hout) {
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, A, sam, sak, B, sbk, sbn, C, ldc, M, N,
                     K, beta, bias, Hout);
}

int ensure_capacity(spx_surrogate* s, int64_t rows) {
  if (rows <= s->cap) return SPX_OK;
  float** bufs[] = {&s->X, &s->T, &s->Z1, &s->H1, &s->Z2, &s->H2, &s->Y, &s->dY, &s->dH, &s->dZ};
  for (float** b : bufs)
    if (*b) { (void)hipFree(*b); *b = nullptr; }
  const int64_t H = s->H;
  const int64_t sizes[] = {rows * s->n_in, rows * s->n_out, rows * H, rows * H, rows * H,
                           rows * H,       rows * s->n_out, rows * s->n_out, rows * H, rows * H};
  for (int i = 0; i < 10; ++i)
    if (hipMalloc(bufs[i], sizeof(float) * std::max<int64_t>(1, sizes[i])) != hipSuccess)
      return fail(SPX_ERR_MEMORY_ALLOC, "surrogate activation buffers");
  s->cap = rows;
  return SPX_OK;
}

// Forward pass of `rows` standardised rows already in s->X; output in s->Y.
void forward(spx_surrogate* s, int64_t rows) {
  hipStream_t st = s->h->stream;
  const int H = s->H;
  const float* w = s->d_w;
  gemm(st, s->X, s->n_in, 1, w + s->off_W1(), H, 1, s->Z1, H, (int)rows, H, s->n_in, 0.f,
       w + s->off_b1(), s->H1);
  gemm(st, s->H1, H, 1, w + s->off_W2(), H, 1, s->Z2, H, (int)rows, H, H, 0.f, w + s->off_b2(),
       s->H2);
  gemm(st, s->H2, H, 1, w + s->off_W3(), s->n_out, 1, s->Y, s->n_out, (int)rows, s->n_out, H, 0.f,
       w + s->off_b3(), nullptr);
}

// Loss + gradient of the standardised MSE into s->d_grad (targets in s->T);
// `wout` weights the output columns (null = unweighted).
void backward(spx_surrogate* s, int64_t rows, const float* wout) {
  hipStream_t st = s->h->stream;
  const int H = s->H, ni = s->n_in, no = s->n_out;
  const float* w = s->d_w;
  float* g = s->d_grad;
  const int64_t cy = rows * no, ch = rows * H;
  hipLaunchKernelGGL(k_mse, dim3(blocks(cy)), dim3(256), 0, st, s->Y, s->T, cy, no, wout, s->dY,
                     s->d_loss);
  // layer 3: dW3 = H2^T dY, db3 = colsum dY, dH2 = dY W3^T
  gemm(st, s->H2, 1, H, s->dY, no, 1, g + s->off_W3(), no, H, no, (int)rows, 0.f, nullptr,
       nullptr);
  hipLaunchKernelGGL(k_colsum, dim3(no), dim3(256), 0, st, s->dY, rows, no, g + s->off_b3());
  gemm(st, s->dY, no, 1, w + s->off_W3(), 1, no, s->dH, H, (int)rows, H, no, 0.f, nullptr, nullptr);
  hipLaunchKernelGGL(k_silu_backward, dim3(blocks(ch)), dim3(256), 0, st, s->dH, s->Z2, ch, s->dZ);
  // layer 2
  gemm(st, s->H1, 1, H, s->dZ, H, 1, g + s->off_W2(), H, H, H, (int)rows, 0.f, nullptr, nullptr);
  hipLaunchKernelGGL(k_colsum, dim3(H), dim3(256), 0, st, s->dZ, rows, H, g + s->off_b2());
  gemm(st, s->dZ, H, 1, w + s->off_W2(), 1, H, s->dH, H, (int)rows, H, H, 0.f, nullptr, nullptr);
  hipLaunchKernelGGL(k_silu_backward, dim3(blocks(ch)), dim3(256), 0, st, s->dH, s->Z1, ch, s->dZ);
  // layer 1
  gemm(st, s->X, 1, ni, s->dZ, H, 1, g + s->off_W1(), H, ni, H, (int)rows, 0.f, nullptr, nullptr);
  hipLaunchKernelGGL(k_colsum, dim3(H), dim3(256), 0, st, s->dZ, rows, H, g + s->off_b1());
}

}  // namespace

int surrogate_predict_locked(spx_surrogate* s, const double* d_params, int64_t B, int32_t n_x,
                             int32_t P, double* d_x0) {
  if (s->slots.empty()) return fail(SPX_ERR_INVALID_ARGUMENT, "surrogate inputs not set");
  if (P != s->P) return fail(SPX_ERR_INVALID_ARGUMENT, "surrogate was set up for another P");
  if (n_x < s->n_out) return fail(SPX_ERR_INVALID_DIMENSION, "n_x smaller than surrogate output");
  hipStream_t st = s->h->stream;
  // fused path: hidden width 64..256 in steps of 64, inputs and outputs <= 64
  const int H = s->H, kp = (s->n_in + 15) / 16 * 16;
  if (H % 64 == 0 && H <= 256 && s->n_in <= 64 && s->n_out <= 64) {
    const int64_t n1 = (int64_t)H * kp, n2 = (int64_t)H * H, n3 = (int64_t)64 * H;
    if (s->wt_cap < n1 + n2 + n3) {
      if (s->d_wt) (void)hipFree(s->d_wt);
      s->d_wt = nullptr;
      s->wt_cap = 0;
      if (hipMalloc(&s->d_wt, sizeof(unsigned short) * (n1 + n2 + n3)) != hipSuccess)
        return fail(SPX_ERR_MEMORY_ALLOC, "surrogate BF16 weights");
      s->wt_cap = n1 + n2 + n3;
    }
    unsigned short *w1t = s->d_wt, *w2t = w1t + n1, *w3t = w2t + n2;
    const float* w = s->d_w;
    if (s->wt_version != s->w_version) {
      // the output layer's columns past n_out are zero rows of W3t
      hipLaunchKernelGGL(k_weights_to_bf16t, dim3(blocks(n1)), dim3(256), 0, st, w + s->off_W1(),
                         s->n_in, H, kp, H, w1t);
      hipLaunchKernelGGL(k_weights_to_bf16t, dim3(blocks(n2)), dim3(256), 0, st, w + s->off_W2(),
                         H, H, H, H, w2t);
      hipLaunchKernelGGL(k_weights_to_bf16t, dim3(blocks(n3)), dim3(256), 0, st, w + s->off_W3(),
                         H, s->n_out, H, 64, w3t);
      SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
      s->wt_version = s->w_version;
    }
    const size_t sm = sizeof(unsigned short) * kInfRows * ((kp + 8) + 2 * (H + 8));
    const int64_t nb = std::min<int64_t>((B + kInfRows - 1) / kInfRows, 1 << 20);
    auto launch = [&](auto kern) -> hipError_t {
      hipError_t e = hipFuncSetAttribute((const void*)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), sm, st, d_params, P, s->d_slots,
                         s->n_in, s->d_norm, w1t, w + s->off_b1(), w2t, w + s->off_b2(), w3t,
                         w + s->off_b3(), s->n_out, n_x, B, d_x0);
      return hipGetLastError();
    };
    auto by_k = [&](auto ht) -> hipError_t {
      constexpr int HT = decltype(ht)::value;
      return kp == 16 ? launch(k_mlp_infer<HT, 16>) : kp == 32 ? launch(k_mlp_infer<HT, 32>)
           : kp == 48 ? launch(k_mlp_infer<HT, 48>) : launch(k_mlp_infer<HT, 64>);
    };
    hipError_t e = H == 64 ? by_k(std::integral_constant<int, 1>{})
                 : H == 128 ? by_k(std::integral_constant<int, 2>{})
                 : H == 192 ? by_k(std::integral_constant<int, 3>{})
                 : by_k(std::integral_constant<int, 4>{});
    SPX_HIP(e, SPX_ERR_BACKEND);
    return SPX_OK;
  }
  const int64_t chunk = 65536;
  int rc = ensure_capacity(s, std::min<int64_t>(B, chunk));
  if (rc) return rc;
  for (int64_t base = 0; base < B; base += chunk) {
    const int64_t rows = std::min<int64_t>(chunk, B - base);
    hipLaunchKernelGGL(k_gather_inputs, dim3(blocks(rows * s->n_in)), dim3(256), 0, st, d_params,
                       P, s->d_slots, s->n_in, s->d_norm, nullptr, base, rows, s->X);
    forward(s, rows);
    hipLaunchKernelGGL(k_write_prediction, dim3(blocks(rows * n_x)), dim3(256), 0, st, s->Y,
                       s->n_out, n_x, s->n_in, s->d_norm, base, rows, d_x0);
  }
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  return SPX_OK;
}

}  // namespace spx

// ===================================================================== C ABI
extern "C" {

int spx_surrogate_create(spx_handle* h, int32_t n_in, int32_t n_hidden, int32_t n_out,
                         uint64_t seed, spx_surrogate** out) {
  if (!h || !out) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle or output");
  if (n_in < 1 || n_hidden < 1 || n_out < 1 || n_in > 4096 || n_hidden > 4096 || n_out > 4096)
    return fail(SPX_ERR_INVALID_DIMENSION, "surrogate layer sizes must be 1..4096");
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  auto* s = new spx_surrogate;
  s->h = h;
  s->n_in = n_in;
  s->H = n_hidden;
  s->n_out = n_out;
  s->nw = (int64_t)n_in * n_hidden + n_hidden + (int64_t)n_hidden * n_hidden + n_hidden +
          (int64_t)n_hidden * n_out + n_out;
  // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases.
  std::vector<float> w(s->nw, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int64_t off, int64_t count, int fan_in) {
    std::uniform_real_distribution<float> U(-1.f / std::sqrt((float)fan_in),
                                            1.f / std::sqrt((float)fan_in));
    for (int64_t i = 0; i < count; ++i) w[off + i] = U(rng);
  };
  fill(s->off_W1(), (int64_t)n_in * n_hidden, n_in);
  fill(s->off_W2(), (int64_t)n_hidden * n_hidden, n_hidden);
  fill(s->off_W3(), (int64_t)n_hidden * n_out, n_hidden);
  std::vector<float> norm(2 * n_in + 2 * n_out, 0.f);
  for (int i = 0; i < n_in; ++i) norm[n_in + i] = 1.f;
  for (int j = 0; j < n_out; ++j) norm[2 * n_in + n_out + j] = 1.f;
  bool ok = hipMalloc(&s->d_w, sizeof(float) * s->nw) == hipSuccess &&
            hipMalloc(&s->d_grad, sizeof(float) * s->nw) == hipSuccess &&
            hipMalloc(&s->d_m, sizeof(float) * s->nw) == hipSuccess &&
            hipMalloc(&s->d_v, sizeof(float) * s->nw) == hipSuccess &&
            hipMalloc(&s->d_norm, sizeof(float) * norm.size()) == hipSuccess &&
            hipMalloc(&s->d_loss, sizeof(float)) == hipSuccess;
  if (!ok) {
    void* ptrs[] = {s->d_w, s->d_grad, s->d_m, s->d_v, s->d_norm, s->d_loss};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    delete s;
    return fail(SPX_ERR_MEMORY_ALLOC, "surrogate weights");
  }
  ok = hipMemcpy(s->d_w, w.data(), sizeof(float) * s->nw, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemset(s->d_m, 0, sizeof(float) * s->nw) == hipSuccess &&
       hipMemset(s->d_v, 0, sizeof(float) * s->nw) == hipSuccess &&
       hipMemcpy(s->d_norm, norm.data(), sizeof(float) * norm.size(), hipMemcpyHostToDevice) ==
           hipSuccess;
  if (!ok) {  // the handle's mutex is held here, so free directly
    void* ptrs[] = {s->d_w, s->d_grad, s->d_m, s->d_v, s->d_norm, s->d_loss};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    delete s;
    return fail(SPX_ERR_TRANSFER, "surrogate initial upload");
  }
  *out = s;
  return SPX_OK;
}

int spx_surrogate_destroy(spx_surrogate* s) {
  if (!s) return SPX_OK;
  spx_handle* h = s->h;
  {
    std::lock_guard<std::mutex> lk(h->mu);
    DeviceGuard dg(h->device);
    (void)hipStreamSynchronize(h->stream);
    if (h->surrogate == s) h->surrogate = nullptr;
    void* ptrs[] = {s->d_w, s->d_grad, s->d_m, s->d_v, s->d_norm, s->d_loss, s->d_slots,
                    s->d_perm, s->d_wout, s->X, s->T, s->Z1, s->H1, s->Z2, s->H2, s->Y, s->dY, s->dH, s->dZ,
                    s->d_wt};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  }
  delete s;
  return SPX_OK;
}

int spx_surrogate_set_inputs(spx_surrogate* s, const int32_t* slots, int32_t P) {
  if (!s || !slots) return fail(SPX_ERR_INVALID_ARGUMENT, "null argument");
  for (int i = 0; i < s->n_in; ++i)
    if (slots[i] < 0 || slots[i] >= P)
      return fail(SPX_ERR_INVALID_ARGUMENT, "input slot out of range");
  std::lock_guard<std::mutex> lk(s->h->mu);
  DeviceGuard dg(s->h->device);
  if (!s->d_slots && hipMalloc(&s->d_slots, sizeof(int) * s->n_in) != hipSuccess)
    return fail(SPX_ERR_MEMORY_ALLOC, "surrogate slots");
  s->slots.assign(slots, slots + s->n_in);
  s->P = P;
  SPX_HIP(hipMemcpy(s->d_slots, slots, sizeof(int) * s->n_in, hipMemcpyHostToDevice),
          SPX_ERR_TRANSFER);
  return SPX_OK;
}

int64_t spx_surrogate_num_weights(const spx_surrogate* s) { return s ? s->nw : 0; }

int spx_surrogate_get_weights(spx_surrogate* s, float* host, int64_t count) {
  if (!s || !host || count != s->nw) return fail(SPX_ERR_INVALID_DIMENSION, "weight count");
  std::lock_guard<std::mutex> lk(s->h->mu);
  DeviceGuard dg(s->h->device);
  SPX_HIP(hipStreamSynchronize(s->h->stream), SPX_ERR_BACKEND);
  SPX_HIP(hipMemcpy(host, s->d_w, sizeof(float) * count, hipMemcpyDeviceToHost), SPX_ERR_TRANSFER);
  return SPX_OK;
}

int spx_surrogate_set_weights(spx_surrogate* s, const float* host, int64_t count) {
  if (!s || !host || count != s->nw) return fail(SPX_ERR_INVALID_DIMENSION, "weight count");
  std::lock_guard<std::mutex> lk(s->h->mu);
  DeviceGuard dg(s->h->device);
  SPX_HIP(hipStreamSynchronize(s->h->stream), SPX_ERR_BACKEND);
  SPX_HIP(hipMemcpy(s->d_w, host, sizeof(float) * count, hipMemcpyHostToDevice), SPX_ERR_TRANSFER);
  ++s->w_version;
  return SPX_OK;
}

int spx_surrogate_set_normalization(spx_surrogate* s, const float* in_mean, const float* in_std,
                                    const float* out_mean, const float* out_std) {
  if (!s || !in_mean || !in_std || !out_mean || !out_std)
    return fail(SPX_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<float> norm;
  norm.insert(norm.end(), in_mean, in_mean + s->n_in);
  norm.insert(norm.end(), in_std, in_std + s->n_in);
  norm.insert(norm.end(), out_mean, out_mean + s->n_out);
  norm.insert(norm.end(), out_std, out_std + s->n_out);
  for (float v : norm)
    if (!std::isfinite(v)) return fail(SPX_ERR_INVALID_ARGUMENT, "non-finite normalisation");
  std::lock_guard<std::mutex> lk(s->h->mu);
  DeviceGuard dg(s->h->device);
  SPX_HIP(hipMemcpy(s->d_norm, norm.data(), sizeof(float) * norm.size(), hipMemcpyHostToDevice),
          SPX_ERR_TRANSFER);
  return SPX_OK;
}

int spx_surrogate_train_device(spx_surrogate* s, const double* d_params, const double* d_x,
                               int32_t n_x, const uint8_t* d_mask, int64_t B, int32_t epochs,
                               int32_t batch, float lr, float tol_rel, float tol_abs,
                               uint64_t shuffle_seed, float* loss_history) {
  if (!s) return fail(SPX_ERR_INVALID_ARGUMENT, "null surrogate");
  if (s->slots.empty()) return fail(SPX_ERR_INVALID_ARGUMENT, "surrogate inputs not set");
  if (n_x < s->n_out) return fail(SPX_ERR_INVALID_DIMENSION, "n_x smaller than surrogate output");
  if (B < 1 || epochs < 0 || batch < 1 || !(lr > 0.f))
    return fail(SPX_ERR_INVALID_ARGUMENT, "bad training arguments");
  spx_handle* h = s->h;
  std::lock_guard<std::mutex> lk(h->mu);
  DeviceGuard dg(h->device);
  hipStream_t st = h->stream;
  int rc = ensure_capacity(s, std::min<int64_t>(batch, B));
  if (rc) return rc;
  // normalisation from the (masked) training rows
  std::vector<int> idx(s->n_out);
  for (int j = 0; j < s->n_out; ++j) idx[j] = j;
  int* d_oidx = nullptr;
  SPX_HIP(hipMalloc(&d_oidx, sizeof(int) * s->n_out), SPX_ERR_MEMORY_ALLOC);
  SPX_HIP(hipMemcpy(d_oidx, idx.data(), sizeof(int) * s->n_out, hipMemcpyHostToDevice),
          SPX_ERR_TRANSFER);
  float* nm = s->d_norm;
  hipLaunchKernelGGL(k_colstats, dim3(s->n_in), dim3(256), 0, st, d_params, (int64_t)s->P,
                     s->d_slots, B, d_mask, nm, nm + s->n_in);
  hipLaunchKernelGGL(k_colstats, dim3(s->n_out), dim3(256), 0, st, d_x, (int64_t)n_x, d_oidx, B,
                     d_mask, nm + 2 * s->n_in, nm + 2 * s->n_in + s->n_out);
  // Output weights: a warm start is good when every node lands inside its
  // Newton tolerance tol_j = tol_rel |mean_j| + tol_abs, so errors are
  // weighted by (std_j / tol_j)^2, normalised to mean 1 and clipped to 1e-3..1e3.
  const bool weighted = tol_rel > 0.f || tol_abs > 0.f;
  if (weighted) {
    std::vector<float> norm(2 * s->n_in + 2 * s->n_out);
    SPX_HIP(hipMemcpyAsync(norm.data(), nm, sizeof(float) * norm.size(), hipMemcpyDeviceToHost, st),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(st), SPX_ERR_TRANSFER);
    std::vector<float> w(s->n_out);
    double sum = 0;
    for (int j = 0; j < s->n_out; ++j) {
      const double mu = norm[2 * s->n_in + j], sd = norm[2 * s->n_in + s->n_out + j];
      const double tol = tol_rel * std::fabs(mu) + tol_abs;
      w[j] = (float)std::pow(sd / std::max(tol, 1e-30), 2.0);
      sum += w[j];
    }
    for (int j = 0; j < s->n_out; ++j)
      w[j] = std::min(1e3f, std::max(1e-3f, (float)(w[j] * s->n_out / sum)));
    if (!s->d_wout) SPX_HIP(hipMalloc(&s->d_wout, sizeof(float) * s->n_out), SPX_ERR_MEMORY_ALLOC);
    SPX_HIP(hipMemcpyAsync(s->d_wout, w.data(), sizeof(float) * s->n_out, hipMemcpyHostToDevice,
                           st),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(st), SPX_ERR_TRANSFER);
  }
  // rows used for training: the masked ones
  std::vector<int> rows;
  if (d_mask) {
    std::vector<uint8_t> m(B);
    SPX_HIP(hipMemcpyAsync(m.data(), d_mask, B, hipMemcpyDeviceToHost, st), SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(st), SPX_ERR_TRANSFER);
    for (int64_t i = 0; i < B; ++i)
      if (m[i]) rows.push_back((int)i);
  } else {
    rows.resize(B);
    for (int64_t i = 0; i < B; ++i) rows[i] = (int)i;
  }
  (void)hipFree(d_oidx);
  const int64_t R = (int64_t)rows.size();
  if (R == 0) return fail(SPX_ERR_INVALID_ARGUMENT, "no training rows");
  if (s->perm_cap < R) {
    if (s->d_perm) (void)hipFree(s->d_perm);
    SPX_HIP(hipMalloc(&s->d_perm, sizeof(int) * R), SPX_ERR_MEMORY_ALLOC);
    s->perm_cap = R;
  }
  std::mt19937_64 rng(shuffle_seed);
  const int64_t steps_per_epoch = (R + batch - 1) / batch;
  const int64_t total = steps_per_epoch * std::max(1, epochs);
  int64_t step = 0;
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(rows.begin(), rows.end(), rng);
    SPX_HIP(hipMemcpyAsync(s->d_perm, rows.data(), sizeof(int) * R, hipMemcpyHostToDevice, st),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipMemsetAsync(s->d_loss, 0, sizeof(float), st), SPX_ERR_BACKEND);
    for (int64_t b0 = 0; b0 < R; b0 += batch, ++step) {
      const int64_t nb = std::min<int64_t>(batch, R - b0);
      hipLaunchKernelGGL(k_gather_inputs, dim3(blocks(nb * s->n_in)), dim3(256), 0, st, d_params,
                         s->P, s->d_slots, s->n_in, s->d_norm, s->d_perm, b0, nb, s->X);
      hipLaunchKernelGGL(k_gather_targets, dim3(blocks(nb * s->n_out)), dim3(256), 0, st, d_x,
                         n_x, s->n_out, s->n_in, s->d_norm, s->d_perm, b0, nb, s->T);
      forward(s, nb);
      backward(s, nb, weighted ? s->d_wout : nullptr);
      // cosine decay to 5% of lr, Adam with bias correction
      const double frac = (double)step / (double)std::max<int64_t>(1, total - 1);
      const float lrt = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac))));
      ++s->adam_t;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)s->adam_t), c2 = 1.f - std::pow(b2, (float)s->adam_t);
      ++s->w_version;
      hipLaunchKernelGGL(k_adam, dim3(blocks(s->nw)), dim3(256), 0, st, s->d_w, s->d_grad, s->d_m,
                         s->d_v, s->nw, lrt, b1, b2, 1e-8f, c1, c2);
    }
    float l = 0.f;
    SPX_HIP(hipMemcpyAsync(&l, s->d_loss, sizeof(float), hipMemcpyDeviceToHost, st),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(st), SPX_ERR_BACKEND);
    if (loss_history) loss_history[e] = l / (float)steps_per_epoch;
  }
  SPX_HIP(hipStreamSynchronize(st), SPX_ERR_BACKEND);
  return SPX_OK;
}

int spx_surrogate_predict_device(spx_surrogate* s, const double* d_params, int64_t B,
                                 int32_t n_x, double* d_x0) {
  if (!s || !d_params || !d_x0) return fail(SPX_ERR_INVALID_ARGUMENT, "null argument");
  if (B < 0) return fail(SPX_ERR_INVALID_DIMENSION, "negative batch");
  if (B == 0) return SPX_OK;
  std::lock_guard<std::mutex> lk(s->h->mu);
  DeviceGuard dg(s->h->device);
  return surrogate_predict_locked(s, d_params, B, n_x, s->P, d_x0);
}

int spx_surrogate_gradient_device(spx_surrogate* s, const float* d_in, const float* d_target,
                                  int64_t B, float* d_grad, float* loss) {
  if (!s || !d_in || !d_target || !d_grad) return fail(SPX_ERR_INVALID_ARGUMENT, "null argument");
  if (B < 1) return fail(SPX_ERR_INVALID_DIMENSION, "empty batch");
  std::lock_guard<std::mutex> lk(s->h->mu);
  DeviceGuard dg(s->h->device);
  hipStream_t st = s->h->stream;
  int rc = ensure_capacity(s, B);
  if (rc) return rc;
  SPX_HIP(hipMemcpyAsync(s->X, d_in, sizeof(float) * B * s->n_in, hipMemcpyDeviceToDevice, st),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipMemcpyAsync(s->T, d_target, sizeof(float) * B * s->n_out, hipMemcpyDeviceToDevice,
                         st),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipMemsetAsync(s->d_loss, 0, sizeof(float), st), SPX_ERR_BACKEND);
  forward(s, B);
  backward(s, B, nullptr);
  SPX_HIP(hipMemcpyAsync(d_grad, s->d_grad, sizeof(float) * s->nw, hipMemcpyDeviceToDevice, st),
          SPX_ERR_TRANSFER);
  float l = 0.f;
  SPX_HIP(hipMemcpyAsync(&l, s->d_loss, sizeof(float), hipMemcpyDeviceToHost, st),
          SPX_ERR_TRANSFER);
  SPX_HIP(hipStreamSynchronize(st), SPX_ERR_BACKEND);
  if (loss) *loss = l;
  return SPX_OK;
}

int spx_set_surrogate(spx_handle* h, spx_surrogate* s) {
  if (!h) return fail(SPX_ERR_INVALID_ARGUMENT, "null handle");
  if (s && s->h != h) return fail(SPX_ERR_INVALID_ARGUMENT, "surrogate belongs to another handle");
  std::lock_guard<std::mutex> lk(h->mu);
  h->surrogate = s;
  return SPX_OK;
}

}  // extern "C"
