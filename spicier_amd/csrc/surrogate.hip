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
  // fused training (32 -> 256 -> 256 -> 64): BF16 copies of W2 and W3 in
  // their own [in][out] layout (the backward products read them so), the
  // row pass's transposed activations, and its per-workgroup bias partials
  unsigned short* d_wn = nullptr;
  unsigned short* d_act = nullptr;
  int64_t act_cap = 0;  // rows
  float* d_psum = nullptr;
  int64_t psum_cap = 0;  // workgroups
  float* d_partial = nullptr;  // split-K partial tiles of the weight gradients
  float* d_gx = nullptr;       // an epoch's inputs and targets in shuffled order
  int64_t gx_cap = 0;          // rows
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

// Round to nearest even: gfx950's v_cvt_pk_bf16_f32 (one instruction, the
// same bits as adding 0x7fff + lsb and truncating, for every finite value).
__device__ __forceinline__ unsigned short to_bf16(float v) {
  return __builtin_bit_cast(unsigned short, (__bf16)v);
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
    // inputs (k_gather_inputs' arithmetic), padded columns zero.  Thread t
    // takes column t % kp of rows t / kp, + 256 / kp, ...: its slot and
    // normalisation are loaded once and its rows' parameter loads go out
    // together (one round trip, not one per element)
    {
      constexpr int RS = 256 / KP, M = (kInfRows + RS - 1) / RS;
      const int i = tid % kp, rr = tid / kp;
      if (rr < RS) {
        const bool in = i < n_in;
        const int sl = in ? slots[i] : 0;
        const float mu = in ? norm[i] : 0.f, sd = in ? norm[n_in + i] : 1.f;
        double pv[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int r = rr + m * RS;
          const int64_t row = r0 + r;
          pv[m] = (in && r < kInfRows && row < B) ? params[row * P + sl] : 0.0;
        }
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int r = rr + m * RS;
          if (r < kInfRows) {
            const int64_t row = r0 + r;
            float v = 0.f;
            if (in && row < B) v = (float)((pv[m] - mu) / sd);
            Xs[r * ldx + i] = to_bf16(v);
          }
        }
      }
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


// ------------------------------------------------------- fused training
// One training step of the 32 -> 256 -> 256 -> 64 network in two launches
// instead of twenty-odd (the layered path above is launch-bound: its
// GEMMs take 16 or 64 workgroups each).
//
// k_train_rows: 32 batch rows per workgroup, four waves.  Inputs and
// targets are gathered and standardised straight from the sweep (or taken
// from FP32 arrays), the forward pass runs on v_mfma_f32_16x16x16_bf16 with
// the BF16 activations in LDS and the pre-activations Z1, Z2 in the waves'
// accumulator registers, then the loss gradient dY and the backward products
// dH2 = dY W3^T, dH1 = dZ2 W2^T with their SiLU derivatives.  It writes the
// activations the weight gradients need (X, H1, H2, dY, dZ2, dZ1) as BF16
// transposed to [feature][row], and per-workgroup column sums of dY, dZ2,
// dZ1 in FP32 for the bias gradients.  Every GEMM operand is rounded to BF16
// and accumulated in FP32, as the layered path and the oracle's BF16 model.
//
// k_train_weights: the output tiles of dW1 = X^T dZ1 (32 x 64 tiles),
// dW2 = H1^T dZ2 and dW3 = H2^T dY (64 x 64), each a K = rows reduction on
// MFMA split over `ksplit` workgroups that store partial tiles.
// k_train_apply: per tile, the partials summed in order, then with `adam`
// the Adam step (k_adam's elementwise update) and the BF16 copies the next
// step and the inference kernel read, otherwise the gradient; one more
// workgroup does the biases (the row pass's partials in workgroup order).
// The split and its reduction are two launches because a device-wide
// release inside one kernel writes back L2 on this GPU, which costs more
// than a launch.
constexpr int kTrRows = 32;
constexpr int kTrThreads = 512;  // eight waves: 32 hidden columns each

struct TrainRowsArgs {
  const double* params;  // sweep inputs through perm (Xin == null)
  int P;
  const int* slots;
  const int* perm;
  int64_t base;
  const double* x;  // sweep targets (Tin == null)
  int n_x;
  const float* Xin;  // given inputs [nb][NI] / targets [nb][NO]
  const float* Tin;
  const float* norm;
  const float* wout;
  int64_t nb;
  float count;  // nb * NO
  const unsigned short *W1t, *W2t, *W3t, *W2n, *W3n;
  const float *b1, *b2, *b3;
  unsigned short *XT, *H1T, *H2T, *dYT, *dZ2T, *dZ1T;
  int64_t ldr;  // row stride of the transposed activations (rows padded to kTrRows)
  float* psum;  // [workgroup][NO + 2H]: column sums of dY, dZ2, dZ1
  float* loss;
};

// acc[i][j] = sum_k A[i*16 + r][k] * Bt[n0 + j*16 + c][k] over K (multiple of
// 16), rows 0..31: A BF16 row-major in LDS (stride lda), Bt BF16 [n][k] in
// global memory (stride ldb).  Lane l holds rows i*16 + 4(l/16) + t, column
// n0 + j*16 + l%16 of the result in acc[i][j][t].
template <int NT, int K>
__device__ __forceinline__ void rows_mm(const unsigned short* A, int lda,
                                        const unsigned short* __restrict__ Bt, int ldb, int n0,
                                        int lane, floatx4 (&acc)[2][NT]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int r16 = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll 8
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 b[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const uint2 w =
          *reinterpret_cast<const uint2*>(Bt + (int64_t)(n0 + j * 16 + r16) * ldb + k0 + kq);
      b[j][0] = (short)(w.x & 0xffffu);
      b[j][1] = (short)(w.x >> 16);
      b[j][2] = (short)(w.y & 0xffffu);
      b[j][3] = (short)(w.y >> 16);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint2 av = *reinterpret_cast<const uint2*>(A + (i * 16 + r16) * lda + k0 + kq);
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
}

// The first P k-steps of a rows_mm B operand, loaded ahead (before the
// barrier that precedes the layer), so the layer starts on registers rather
// than on a round trip to memory.
template <int NT, int P>
struct BPre {
  uint2 w[P > 0 ? P : 1][NT];
};
template <int NT, int P>
__device__ __forceinline__ void prefetch_b(const unsigned short* __restrict__ Bt, int ldb, int n0,
                                           int lane, BPre<NT, P>& pre) {
  const int r16 = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
  for (int q = 0; q < P; ++q)
#pragma unroll
    for (int j = 0; j < NT; ++j)
      pre.w[q][j] =
          *reinterpret_cast<const uint2*>(Bt + (int64_t)(n0 + j * 16 + r16) * ldb + q * 16 + kq);
}

// rows_mm with the first P k-steps of B taken from `pre`
template <int NT, int K, int P>
__device__ __forceinline__ void rows_mm_pre(const unsigned short* A, int lda,
                                            const unsigned short* __restrict__ Bt, int ldb, int n0,
                                            int lane, floatx4 (&acc)[2][NT],
                                            const BPre<NT, P>& pre) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int r16 = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
  for (int q = 0; q < K / 16; ++q) {
    const int k0 = q * 16;
    bf16x4 b[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const uint2 w = q < P ? pre.w[q < P ? q : 0][j]
                            : *reinterpret_cast<const uint2*>(
                                  Bt + (int64_t)(n0 + j * 16 + r16) * ldb + k0 + kq);
      b[j][0] = (short)(w.x & 0xffffu);
      b[j][1] = (short)(w.x >> 16);
      b[j][2] = (short)(w.y & 0xffffu);
      b[j][3] = (short)(w.y >> 16);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint2 av = *reinterpret_cast<const uint2*>(A + (i * 16 + r16) * lda + k0 + kq);
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
}

template <int NT, int P, int K>
__device__ __forceinline__ void rows_mm_pre(const unsigned short* A, int lda,
                                            const unsigned short* __restrict__ Bt, int ldb, int n0,
                                            int lane, floatx4 (&acc)[2][NT],
                                            const BPre<NT, P>& pre, std::integral_constant<int, K>) {
  rows_mm_pre<NT, K, P>(A, lda, Bt, ldb, n0, lane, acc, pre);
}

// sum over the four lanes l, l^16, l^32, l^48 (one column of a 16-wide tile)
__device__ __forceinline__ float colgroup_sum(float v) {
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

template <int NI, int H, int NO>
__host__ __device__ constexpr size_t train_rows_lds_bytes() {
  return sizeof(unsigned short) * kTrRows * ((NI + 8) + 4 * (H + 8) + (NO + 8));
}

template <int NI, int H, int NO>
__global__ __launch_bounds__(kTrThreads) void k_train_rows(TrainRowsArgs a) {
  static_assert(NI % 16 == 0 && H == 256 && NO == 64, "fused shape");
  constexpr int LX = NI + 8, LH = H + 8, LY = NO + 8;  // BF16 row strides (pad LDS banks)
  extern __shared__ unsigned short lds[];
  unsigned short* Xs = lds;
  unsigned short* H1s = Xs + kTrRows * LX;
  unsigned short* H2s = H1s + kTrRows * LH;
  unsigned short* dYs = H2s + kTrRows * LH;
  unsigned short* dZ2s = dYs + kTrRows * LY;
  unsigned short* dZ1s = dZ2s + kTrRows * LH;
  __shared__ float lsum[kTrThreads / 64];
  __shared__ float Ts[kTrRows][NO];  // targets, gathered with the inputs
  __shared__ float bs[2 * H + 2 * NO];  // b1, b2, b3 and the output weights, loaded up front
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int c16 = lane & 15, rq = (lane >> 4) * 4;
  const int64_t r0 = (int64_t)blockIdx.x * kTrRows;
  const float* om = a.norm + 2 * NI;
  const float* os = om + NO;
  float* ps = a.psum + (int64_t)blockIdx.x * (NO + 2 * H);
  auto src_row = [&](int64_t row) { return a.perm ? (int64_t)a.perm[a.base + row] : a.base + row; };
  // inputs, standardised as k_gather_inputs; rows past nb are zero
  for (int e = tid; e < kTrRows * NI; e += kTrThreads) {
    const int r = e / NI, i = e - r * NI;
    const int64_t row = r0 + r;
    float v = 0.f;
    if (row < a.nb)
      v = a.Xin ? a.Xin[row * NI + i]
                : (float)((a.params[src_row(row) * a.P + a.slots[i]] - a.norm[i]) / a.norm[NI + i]);
    Xs[r * LX + i] = to_bf16(v);
  }
  // biases and output weights: their loads overlap the inputs' instead of
  // sitting on each layer's critical path
  for (int c = tid; c < 2 * H + 2 * NO; c += kTrThreads)
    bs[c] = c < H ? a.b1[c] : c < 2 * H ? a.b2[c - H] : c < 2 * H + NO ? a.b3[c - 2 * H]
          : (a.wout ? a.wout[c - 2 * H - NO] : 1.f);
  // targets, standardised as k_gather_targets (their loads overlap the inputs')
  for (int e = tid; e < kTrRows * NO; e += kTrThreads) {
    const int r = e / NO, c = e - r * NO;
    const int64_t row = r0 + r;
    float tv = 0.f;
    if (row < a.nb)
      tv = a.Tin ? a.Tin[row * NO + c] : (float)((a.x[src_row(row) * a.n_x + c] - om[c]) / os[c]);
    Ts[r][c] = tv;
  }
  constexpr int CW = H / (kTrThreads / 64), NT = CW / 16;  // columns, tiles per wave
  // every layer's first B operands, loaded now (they depend on nothing)
  BPre<NT, NI / 16> p1;
  BPre<NT, 4> p2, p5;
  BPre<1, 4> p3;
  BPre<NT, NO / 16> p4;
  prefetch_b(a.W1t, NI, wave * CW, lane, p1);
  prefetch_b(a.W2t, H, wave * CW, lane, p2);
  if (wave < NO / 16) prefetch_b(a.W3t, H, wave * 16, lane, p3);
  prefetch_b(a.W3n, NO, wave * CW, lane, p4);
  prefetch_b(a.W2n, H, wave * CW, lane, p5);
  __syncthreads();
  // forward: Z1, Z2 stay in registers for the SiLU derivatives
  floatx4 z1[2][NT], z2[2][NT];
  rows_mm_pre<NT, NI, NI / 16>(Xs, LX, a.W1t, NI, wave * CW, lane, z1, p1);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int c = wave * CW + j * 16 + c16;
        const float v = z1[i][j][t] + bs[c];
        z1[i][j][t] = v;
        H1s[(i * 16 + rq + t) * LH + c] = to_bf16(v * sigmoidf(v));
      }
  __syncthreads();
  rows_mm_pre<NT, H, 4>(H1s, LH, a.W2t, H, wave * CW, lane, z2, p2);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int c = wave * CW + j * 16 + c16;
        const float v = z2[i][j][t] + bs[H + c];
        z2[i][j][t] = v;
        H2s[(i * 16 + rq + t) * LH + c] = to_bf16(v * sigmoidf(v));
      }
  __syncthreads();
  // output layer and the (weighted) MSE gradient, as k_mse: one 16-column
  // tile per wave for the first NO / 16 waves
  if (wave < NO / 16) {
    floatx4 y[2][1];
    rows_mm_pre<1, H, 4>(H2s, LH, a.W3t, H, wave * 16, lane, y, p3);
    const int c = wave * 16 + c16;
    const float w = bs[2 * H + NO + c];
    float lpart = 0.f, cs = 0.f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = i * 16 + rq + t;
        const int64_t row = r0 + r;
        float dy = 0.f;
        if (row < a.nb) {
          const float d = (y[i][0][t] + bs[2 * H + c]) - Ts[r][c];
          dy = 2.f * w * d / a.count;
          lpart += w * d * d / a.count;
        }
        dYs[r * LY + c] = to_bf16(dy);
        cs += dy;
      }
    cs = colgroup_sum(cs);
    if (lane < 16) ps[c] = cs;
    for (int off = 32; off >= 1; off >>= 1) lpart += __shfl_xor(lpart, off, 64);
    if (lane == 0) lsum[wave] = lpart;
  }
  __syncthreads();
  if (tid == 0) {
    float l = 0.f;
    for (int w = 0; w < NO / 16; ++w) l += lsum[w];
    atomicAdd(a.loss, l);
  }
  // backward: dZ2 = (dY W3^T) * silu'(Z2), dZ1 = (dZ2 W2^T) * silu'(Z1)
  auto backward_layer = [&](const unsigned short* D, int ld, const unsigned short* Wn, int ldw,
                            auto kdim, floatx4 (&z)[2][NT], unsigned short* out, float* pcol,
                            const auto& pre) {
    constexpr int KD = decltype(kdim)::value;
    floatx4 g[2][NT];
    rows_mm_pre(D, ld, Wn, ldw, wave * CW, lane, g, pre, std::integral_constant<int, KD>{});
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int c = wave * CW + j * 16 + c16;
      float cs = 0.f;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const float zz = z[i][j][t], sg = sigmoidf(zz);
          const float dz = g[i][j][t] * sg * (1.f + zz * (1.f - sg));
          out[(i * 16 + rq + t) * LH + c] = to_bf16(dz);
          cs += dz;
        }
      cs = colgroup_sum(cs);
      if (lane < 16) pcol[c] = cs;
    }
  };
  backward_layer(dYs, LY, a.W3n, NO, std::integral_constant<int, NO>{}, z2, dZ2s, ps + NO, p4);
  __syncthreads();
  backward_layer(dZ2s, LH, a.W2n, H, std::integral_constant<int, H>{}, z1, dZ1s, ps + NO + H, p5);
  __syncthreads();
  // the activations the weight gradients read, transposed to [feature][row]
  auto store_t = [&](const unsigned short* S, int ls, int F, unsigned short* D) {
    for (int e = tid; e < F * (kTrRows / 8); e += kTrThreads) {
      const int f = e / (kTrRows / 8), q = e - f * (kTrRows / 8);
      unsigned short v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = S[(q * 8 + t) * ls + f];
      uint4 pk;
      pk.x = v[0] | ((uint32_t)v[1] << 16);
      pk.y = v[2] | ((uint32_t)v[3] << 16);
      pk.z = v[4] | ((uint32_t)v[5] << 16);
      pk.w = v[6] | ((uint32_t)v[7] << 16);
      *reinterpret_cast<uint4*>(D + (int64_t)f * a.ldr + r0 + q * 8) = pk;
    }
  };
  store_t(Xs, LX, NI, a.XT);
  store_t(H1s, LH, H, a.H1T);
  store_t(H2s, LH, H, a.H2T);
  store_t(dYs, LY, NO, a.dYT);
  store_t(dZ2s, LH, H, a.dZ2T);
  store_t(dZ1s, LH, H, a.dZ1T);
}

struct TrainWeightsArgs {
  const unsigned short *XT, *H1T, *H2T, *dYT, *dZ2T, *dZ1T;
  int64_t ldr;  // rows, a multiple of kTrRows
  int row_blocks;
  int ksplit;       // workgroups per tile, each reducing a contiguous share of the rows
  float* partial;   // [tile][ksplit][64 * 64] partial tiles
  const float* psum;
  float* w;  // FP32 master weights W1,b1,W2,b2,W3,b3
  float* m;
  float* v;
  float* grad;  // gradient out (adam == 0)
  unsigned short *W1t, *W2t, *W3t, *W2n, *W3n;
  int adam;
  float lr, b1, b2, eps, c1, c2;
};

__device__ __forceinline__ float adam_update(float* w, float* m, float* v, int64_t i, float g,
                                             const TrainWeightsArgs& a) {
  const float mt = a.b1 * m[i] + (1.f - a.b1) * g;
  const float vt = a.b2 * v[i] + (1.f - a.b2) * g * g;
  m[i] = mt;
  v[i] = vt;
  const float nw = w[i] - a.lr * (mt / a.c1) / (sqrtf(vt / a.c2) + a.eps);
  w[i] = nw;
  return nw;
}

// Which weight matrix a gradient tile belongs to, and where.
struct TrainTile {
  int which;  // 0 dW1 (32-row tiles), 1 dW2, 2 dW3
  int i0, j0, tm, N, ldt;
  int64_t off;
};
template <int NI, int H, int NO>
__device__ __forceinline__ TrainTile train_tile(int b) {
  constexpr int nb1 = (NI / 32) * (H / 64), nb2 = (H / 64) * (H / 64);
  constexpr int64_t ob1 = (int64_t)NI * H, oW2 = ob1 + H, ob2 = oW2 + (int64_t)H * H,
                    oW3 = ob2 + H;
  if (b < nb1) return TrainTile{0, (b / (H / 64)) * 32, (b % (H / 64)) * 64, 32, H, NI, 0};
  if ((b -= nb1) < nb2) return TrainTile{1, (b / (H / 64)) * 64, (b % (H / 64)) * 64, 64, H, H, oW2};
  b -= nb2;
  return TrainTile{2, (b / (NO / 64)) * 64, (b % (NO / 64)) * 64, 64, NO, H, oW3};
}

// Partial weight-gradient tiles: workgroup (tile, ks) reduces its share of
// the rows on MFMA and stores the tile (row-major, 64 columns) in `partial`.
template <int NI, int H, int NO>
__global__ __launch_bounds__(256) void k_train_weights(TrainWeightsArgs a) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int tile_id = blockIdx.x / a.ksplit, ks = blockIdx.x - tile_id * a.ksplit;
  const TrainTile T = train_tile<NI, H, NO>(tile_id);
  const unsigned short* AT = T.which == 0 ? a.XT : T.which == 1 ? a.H1T : a.H2T;
  const unsigned short* DT = T.which == 0 ? a.dZ1T : T.which == 1 ? a.dZ2T : a.dYT;
  const int r16 = lane & 15, kq = (lane >> 4) * 4, c16 = lane & 15, rq = (lane >> 4) * 4;
  const int jb = T.j0 + (wave & 1) * 32;
  float* mine = a.partial + ((int64_t)tile_id * a.ksplit + ks) * 4096;
  const int64_t steps = a.ldr / 16, per = (steps + a.ksplit - 1) / a.ksplit;
  const int64_t kbeg = ks * per * 16, kend = std::min<int64_t>(a.ldr, (ks + 1) * per * 16);
  auto tile = [&](auto ti) {
    constexpr int TI = decltype(ti)::value;  // 16-row tiles per wave
    const int ib = T.i0 + (wave >> 1) * 16 * TI;
    floatx4 acc[TI][2];
#pragma unroll
    for (int ii = 0; ii < TI; ++ii)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) acc[ii][jj] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
    for (int64_t k0 = kbeg; k0 < kend; k0 += 16) {
      bf16x4 av[TI], bv[2];
#pragma unroll
      for (int ii = 0; ii < TI; ++ii) {
        const uint2 u =
            *reinterpret_cast<const uint2*>(AT + (int64_t)(ib + ii * 16 + r16) * a.ldr + k0 + kq);
        av[ii][0] = (short)(u.x & 0xffffu);
        av[ii][1] = (short)(u.x >> 16);
        av[ii][2] = (short)(u.y & 0xffffu);
        av[ii][3] = (short)(u.y >> 16);
      }
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const uint2 u =
            *reinterpret_cast<const uint2*>(DT + (int64_t)(jb + jj * 16 + r16) * a.ldr + k0 + kq);
        bv[jj][0] = (short)(u.x & 0xffffu);
        bv[jj][1] = (short)(u.x >> 16);
        bv[jj][2] = (short)(u.y & 0xffffu);
        bv[jj][3] = (short)(u.y >> 16);
      }
#pragma unroll
      for (int ii = 0; ii < TI; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
          acc[ii][jj] =
              __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[ii], bv[jj], acc[ii][jj], 0, 0, 0);
    }
#pragma unroll
    for (int ii = 0; ii < TI; ++ii)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          mine[(ib + ii * 16 + rq + t - T.i0) * 64 + (jb + jj * 16 + c16 - T.j0)] = acc[ii][jj][t];
  };
  if (T.tm == 64)
    tile(std::integral_constant<int, 2>{});
  else
    tile(std::integral_constant<int, 1>{});
}

// The gradient tiles summed over their shares in order (deterministic), then
// the Adam step with the BF16 copies, or the gradient; the last workgroup
// does the biases from the row pass's partials, in workgroup order.
template <int NI, int H, int NO>
__global__ __launch_bounds__(256) void k_train_apply(TrainWeightsArgs a, int tiles) {
  constexpr int64_t ob1 = (int64_t)NI * H, oW2 = ob1 + H, ob2 = oW2 + (int64_t)H * H,
                    oW3 = ob2 + H, ob3 = oW3 + (int64_t)H * NO;
  constexpr int per_tile = 64 * 64 / 256;  // workgroups per 64 x 64 tile (a 32-row tile uses half)
  const int tid = threadIdx.x;
  if ((int)blockIdx.x >= tiles * per_tile) {  // biases
    const int c = ((int)blockIdx.x - tiles * per_tile) * 256 + tid;
    if (c >= NO + 2 * H) return;
    float g = 0.f;
#pragma unroll 8
    for (int k = 0; k < a.row_blocks; ++k) g += a.psum[(int64_t)k * (NO + 2 * H) + c];
    const int64_t idx = c < NO ? ob3 + c : c < NO + H ? ob2 + (c - NO) : ob1 + (c - NO - H);
    if (a.adam)
      adam_update(a.w, a.m, a.v, idx, g, a);
    else
      a.grad[idx] = g;
    return;
  }
  const int tile = (int)blockIdx.x / per_tile;
  const int q = ((int)blockIdx.x - tile * per_tile) * 256 + tid;  // element of the tile
  const TrainTile T = train_tile<NI, H, NO>(tile);
  if (q >= T.tm * 64) return;
  const float* all = a.partial + (int64_t)tile * a.ksplit * 4096;
  float g = 0.f;
#pragma unroll 8
  for (int k = 0; k < a.ksplit; ++k) g += all[(int64_t)k * 4096 + q];
  const int i = T.i0 + q / 64, j = T.j0 + q % 64;
  const int64_t idx = T.off + (int64_t)i * T.N + j;
  if (a.adam) {
    unsigned short* Wt = T.which == 0 ? a.W1t : T.which == 1 ? a.W2t : a.W3t;
    unsigned short* Wn = T.which == 0 ? nullptr : T.which == 1 ? a.W2n : a.W3n;
    const float nw = adam_update(a.w, a.m, a.v, idx, g, a);
    const unsigned short h = to_bf16(nw);
    Wt[(int64_t)j * T.ldt + i] = h;
    if (Wn) Wn[(int64_t)i * T.N + j] = h;
  } else {
    a.grad[idx] = g;
  }
}

// BF16 copy of n FP32 values in their own layout
__global__ void k_to_bf16(const float* __restrict__ src, int64_t n, unsigned short* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = to_bf16(src[t]);
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

// The fused training kernels' shape (the workload's network).
bool fused_train_shape(const spx_surrogate* s) {
  return s->n_in == 32 && s->H == 256 && s->n_out == 64;
}
constexpr int kFI = 32, kFH = 256, kFO = 64;
// weight-gradient tiles (dW1 32 x 64, dW2 and dW3 64 x 64) and the
// workgroups that share each tile's K = rows reduction
constexpr int kTrTiles = (kFI / 32) * (kFH / 64) + (kFH / 64) * (kFH / 64) + (kFH / 64) * (kFO / 64);
constexpr int kTrSplit = 8;

// BF16 copies of the current FP32 weights for the fused kernels: d_wt (the
// [out][in] copies the inference kernel reads too) and d_wn (W2, W3 as they
// are, [in][out]).
int fused_refresh_bf16(spx_surrogate* s) {
  hipStream_t st = s->h->stream;
  const int64_t n1 = (int64_t)kFH * kFI, n2 = (int64_t)kFH * kFH, n3 = (int64_t)64 * kFH;
  if (s->wt_cap < n1 + n2 + n3) {
    if (s->d_wt) (void)hipFree(s->d_wt);
    s->d_wt = nullptr;
    s->wt_cap = 0;
    if (hipMalloc(&s->d_wt, sizeof(unsigned short) * (n1 + n2 + n3)) != hipSuccess)
      return fail(SPX_ERR_MEMORY_ALLOC, "surrogate BF16 weights");
    s->wt_cap = n1 + n2 + n3;
  }
  if (!s->d_wn && hipMalloc(&s->d_wn, sizeof(unsigned short) * (n2 + (int64_t)kFH * kFO)) != hipSuccess)
    return fail(SPX_ERR_MEMORY_ALLOC, "surrogate BF16 weights");
  unsigned short *w1t = s->d_wt, *w2t = w1t + n1, *w3t = w2t + n2;
  const float* w = s->d_w;
  hipLaunchKernelGGL(k_weights_to_bf16t, dim3(blocks(n1)), dim3(256), 0, st, w + s->off_W1(), kFI,
                     kFH, kFI, kFH, w1t);
  hipLaunchKernelGGL(k_weights_to_bf16t, dim3(blocks(n2)), dim3(256), 0, st, w + s->off_W2(), kFH,
                     kFH, kFH, kFH, w2t);
  hipLaunchKernelGGL(k_weights_to_bf16t, dim3(blocks(n3)), dim3(256), 0, st, w + s->off_W3(), kFH,
                     kFO, kFH, 64, w3t);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n2)), dim3(256), 0, st, w + s->off_W2(), n2, s->d_wn);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks((int64_t)kFH * kFO)), dim3(256), 0, st,
                     w + s->off_W3(), (int64_t)kFH * kFO, s->d_wn + n2);
  SPX_HIP(hipGetLastError(), SPX_ERR_BACKEND);
  s->wt_version = s->w_version;
  return SPX_OK;
}

// Buffers of a fused step over up to `rows` rows.
int fused_capacity(spx_surrogate* s, int64_t rows) {
  const int64_t nwg = (rows + kTrRows - 1) / kTrRows, ldr = nwg * kTrRows;
  if (s->act_cap < ldr) {
    if (s->d_act) (void)hipFree(s->d_act);
    s->d_act = nullptr;
    s->act_cap = 0;
    if (hipMalloc(&s->d_act, sizeof(unsigned short) * ldr * (kFI + 4 * kFH + kFO)) != hipSuccess)
      return fail(SPX_ERR_MEMORY_ALLOC, "surrogate training activations");
    s->act_cap = ldr;
  }
  if (s->psum_cap < nwg) {
    if (s->d_psum) (void)hipFree(s->d_psum);
    s->d_psum = nullptr;
    s->psum_cap = 0;
    if (hipMalloc(&s->d_psum, sizeof(float) * nwg * (kFO + 2 * kFH)) != hipSuccess)
      return fail(SPX_ERR_MEMORY_ALLOC, "surrogate training partial sums");
    s->psum_cap = nwg;
  }
  if (!s->d_partial &&
      hipMalloc(&s->d_partial, sizeof(float) * kTrTiles * kTrSplit * 4096) != hipSuccess)
    return fail(SPX_ERR_MEMORY_ALLOC, "surrogate training partial tiles");
  static bool lds_set = false;  // the attribute is per function, process-wide
  if (!lds_set) {
    SPX_HIP(hipFuncSetAttribute((const void*)k_train_rows<kFI, kFH, kFO>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)train_rows_lds_bytes<kFI, kFH, kFO>()),
            SPX_ERR_BACKEND);
    lds_set = true;
  }
  return SPX_OK;
}

// One fused step over nb rows: the loss gradient of rows taken from the
// sweep (params/x through perm from `base`) or from Xin/Tin, then either an
// Adam step (adam != 0) or the gradient into s->d_grad.
hipError_t fused_step(spx_surrogate* s, const double* params, const double* x, int n_x,
                      const int* perm, int64_t base, const float* Xin, const float* Tin,
                      const float* wout, int64_t nb, int adam, float lr, float c1, float c2) {
  hipStream_t st = s->h->stream;
  const int nwg = (int)((nb + kTrRows - 1) / kTrRows);
  const int64_t ldr = (int64_t)nwg * kTrRows;
  unsigned short* XT = s->d_act;
  unsigned short* H1T = XT + (int64_t)kFI * ldr;
  unsigned short* H2T = H1T + (int64_t)kFH * ldr;
  unsigned short* dYT = H2T + (int64_t)kFH * ldr;
  unsigned short* dZ2T = dYT + (int64_t)kFO * ldr;
  unsigned short* dZ1T = dZ2T + (int64_t)kFH * ldr;
  const int64_t n1 = (int64_t)kFH * kFI, n2 = (int64_t)kFH * kFH;
  unsigned short *w1t = s->d_wt, *w2t = w1t + n1, *w3t = w2t + n2, *w2n = s->d_wn,
                 *w3n = s->d_wn + n2;
  const float* w = s->d_w;
  TrainRowsArgs ra{params, s->P, s->d_slots, perm, base, x, n_x, Xin, Tin, s->d_norm, wout, nb,
                   (float)(nb * kFO), w1t, w2t, w3t, w2n, w3n, w + s->off_b1(), w + s->off_b2(),
                   w + s->off_b3(), XT, H1T, H2T, dYT, dZ2T, dZ1T, ldr, s->d_psum, s->d_loss};
  const size_t lds = train_rows_lds_bytes<kFI, kFH, kFO>();
  hipLaunchKernelGGL((k_train_rows<kFI, kFH, kFO>), dim3(nwg), dim3(kTrThreads), lds, st, ra);
  const int ksplit = (int)std::max<int64_t>(1, std::min<int64_t>(kTrSplit, ldr / 16));
  TrainWeightsArgs wa{XT, H1T, H2T, dYT, dZ2T, dZ1T, ldr, nwg, ksplit, s->d_partial,
                      s->d_psum, s->d_w, s->d_m, s->d_v, s->d_grad, w1t, w2t, w3t, w2n, w3n,
                      adam, lr, 0.9f, 0.999f, 1e-8f, c1, c2};
  hipLaunchKernelGGL((k_train_weights<kFI, kFH, kFO>), dim3(kTrTiles * ksplit), dim3(256), 0, st,
                     wa);
  const int bias_blocks = (kFO + 2 * kFH + 255) / 256;
  hipLaunchKernelGGL((k_train_apply<kFI, kFH, kFO>), dim3(kTrTiles * 16 + bias_blocks), dim3(256),
                     0, st, wa, kTrTiles);
  return hipGetLastError();
}

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
                    s->d_wt, s->d_wn, s->d_act, s->d_psum, s->d_partial, s->d_gx};
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
  const bool fused = fused_train_shape(s);
  if (fused) {
    if ((rc = fused_capacity(s, std::min<int64_t>(batch, R)))) return rc;
    if ((rc = fused_refresh_bf16(s))) return rc;
    if (s->gx_cap < R) {  // each epoch gathers its shuffled rows once, off the steps' path
      if (s->d_gx) (void)hipFree(s->d_gx);
      s->d_gx = nullptr;
      s->gx_cap = 0;
      if (hipMalloc(&s->d_gx, sizeof(float) * R * (s->n_in + s->n_out)) != hipSuccess)
        return fail(SPX_ERR_MEMORY_ALLOC, "surrogate training rows");
      s->gx_cap = R;
    }
  }
  float* gX = fused ? s->d_gx : nullptr;
  float* gT = fused ? s->d_gx + R * s->n_in : nullptr;
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(rows.begin(), rows.end(), rng);
    SPX_HIP(hipMemcpyAsync(s->d_perm, rows.data(), sizeof(int) * R, hipMemcpyHostToDevice, st),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipMemsetAsync(s->d_loss, 0, sizeof(float), st), SPX_ERR_BACKEND);
    if (fused) {
      hipLaunchKernelGGL(k_gather_inputs, dim3(blocks(R * s->n_in)), dim3(256), 0, st, d_params,
                         s->P, s->d_slots, s->n_in, s->d_norm, s->d_perm, (int64_t)0, R, gX);
      hipLaunchKernelGGL(k_gather_targets, dim3(blocks(R * s->n_out)), dim3(256), 0, st, d_x, n_x,
                         s->n_out, s->n_in, s->d_norm, s->d_perm, (int64_t)0, R, gT);
    }
    for (int64_t b0 = 0; b0 < R; b0 += batch, ++step) {
      const int64_t nb = std::min<int64_t>(batch, R - b0);
      // cosine decay to 5% of lr, Adam with bias correction
      const double frac = (double)step / (double)std::max<int64_t>(1, total - 1);
      const float lrt = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * frac))));
      ++s->adam_t;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)s->adam_t), c2 = 1.f - std::pow(b2, (float)s->adam_t);
      ++s->w_version;
      if (fused) {
        SPX_HIP(fused_step(s, nullptr, nullptr, 0, nullptr, 0, gX + b0 * s->n_in,
                           gT + b0 * s->n_out, weighted ? s->d_wout : nullptr, nb, 1, lrt, c1, c2),
                SPX_ERR_BACKEND);
        s->wt_version = s->w_version;  // the step wrote the BF16 copies too
        continue;
      }
      hipLaunchKernelGGL(k_gather_inputs, dim3(blocks(nb * s->n_in)), dim3(256), 0, st, d_params,
                         s->P, s->d_slots, s->n_in, s->d_norm, s->d_perm, b0, nb, s->X);
      hipLaunchKernelGGL(k_gather_targets, dim3(blocks(nb * s->n_out)), dim3(256), 0, st, d_x,
                         n_x, s->n_out, s->n_in, s->d_norm, s->d_perm, b0, nb, s->T);
      forward(s, nb);
      backward(s, nb, weighted ? s->d_wout : nullptr);
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
  int rc;
  if (fused_train_shape(s)) {  // the training step's kernels, without the Adam step
    if ((rc = fused_capacity(s, B)) || (rc = fused_refresh_bf16(s))) return rc;
    SPX_HIP(hipMemsetAsync(s->d_loss, 0, sizeof(float), st), SPX_ERR_BACKEND);
    SPX_HIP(fused_step(s, nullptr, nullptr, 0, nullptr, 0, d_in, d_target, nullptr, B, 0, 0.f,
                       1.f, 1.f),
            SPX_ERR_BACKEND);
    SPX_HIP(hipMemcpyAsync(d_grad, s->d_grad, sizeof(float) * s->nw, hipMemcpyDeviceToDevice, st),
            SPX_ERR_TRANSFER);
    float l = 0.f;
    SPX_HIP(hipMemcpyAsync(&l, s->d_loss, sizeof(float), hipMemcpyDeviceToHost, st),
            SPX_ERR_TRANSFER);
    SPX_HIP(hipStreamSynchronize(st), SPX_ERR_BACKEND);
    if (loss) *loss = l;
    return SPX_OK;
  }
  rc = ensure_capacity(s, B);
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
