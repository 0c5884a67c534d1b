// Learned (k, p, q) proposer for rand_svd: a 34 -> 128 -> 128 -> 3 MLP with
// GELU activations.  Weights and activations are rounded to BF16 and every
// product runs on the BF16 matrix cores (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation; the master weights, biases, losses and Adam state are FP32.
// The reference has no learned component.
//
// Parameters (FP32, packed): W1 [128][34], b1 [128], W2 [128][128], b2 [128],
// W3 [3][128], b3 [3] (row-major, out x in): 21379 values.
//
// One work-group of 256 threads handles 32 samples: it stages the BF16
// weights and its samples in LDS, runs the forward pass and, for training,
// the backward pass, and writes its share of the weight gradient to a
// per-work-group slot.  The slots are summed in a fixed order (deterministic)
// and Adam updates the master weights.
#include "common.h"
#include "kernels.h"

namespace rna {
namespace {

constexpr int MIN = 34, MH = 128, MOUT = 3;
constexpr int KIN = 48;                    // input width padded to the MFMA K step
constexpr int NOUT = 16;                   // output width padded to one MFMA tile
constexpr int SB = 32;                     // samples per work-group
constexpr int MT = 256;                    // threads per work-group
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MH * MIN, OFF_W2 = OFF_B1 + MH,
              OFF_B2 = OFF_W2 + MH * MH, OFF_W3 = OFF_B2 + MH, OFF_B3 = OFF_W3 + MOUT * MH,
              NPARAM = OFF_B3 + MOUT;
static_assert(NPARAM == 21379, "parameter count");

typedef short s16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf(float x) { return (uint16_t)bf16_bits(x); }
__device__ __forceinline__ float fb(uint16_t h) { return bf16_to_f32(h); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C (M x N) = A (M x K) B (K x N) on 16x16x16 BF16 MFMA tiles; M, N, K are
// multiples of 16.  a(i, k) and b(k, j) return BF16 bits; store(i, j, v).
// Lane l of a tile supplies A[l % 16][4 (l / 16) + 0..3] and
// B[4 (l / 16) + 0..3][l % 16]; accumulator v holds C[4 (l / 16) + v][l % 16].
template <class FA, class FB, class FC>
__device__ void wg_mm(int M, int N, int K, FA a, FB b, FC store) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int tn = N / 16, tiles = (M / 16) * tn;
  for (int t = w; t < tiles; t += MT / 64) {
    const int i0 = (t / tn) * 16, j0 = (t % tn) * 16;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
      const int kb = k0 + 4 * (l >> 4);
      s16x4 av, bv;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        av[q] = (short)a(i0 + (l & 15), kb + q);
        bv[q] = (short)b(kb + q, j0 + (l & 15));
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) store(i0 + 4 * (l >> 4) + v, j0 + (l & 15), acc[v]);
  }
}

struct MlpLds {
  uint16_t W1[MH][KIN];        // bf16, input columns >= 34 zero
  uint16_t W2[MH][MH];
  uint16_t W3[NOUT][MH];       // rows >= 3 zero
  uint16_t X[SB][KIN];
  uint16_t H1[SB][MH];         // GELU(Z1) in bf16; later dZ1
  uint16_t H2[SB][MH];         // GELU(Z2) in bf16; later dZ2
  float G1[SB][MH];            // GELU'(Z1)
  float G2[SB][MH];            // GELU'(Z2)
  float Y[SB][NOUT];
  uint16_t DY[SB][NOUT];
  float red[MT / 64];
};

__device__ void stage_weights(const float* P, MlpLds& S) {
  for (int i = threadIdx.x; i < MH * KIN; i += MT) {
    const int o = i / KIN, c = i % KIN;
    S.W1[o][c] = c < MIN ? bf(P[OFF_W1 + o * MIN + c]) : 0;
  }
  for (int i = threadIdx.x; i < MH * MH; i += MT) S.W2[i / MH][i % MH] = bf(P[OFF_W2 + i]);
  for (int i = threadIdx.x; i < NOUT * MH; i += MT) {
    const int o = i / MH, c = i % MH;
    S.W3[o][c] = o < MOUT ? bf(P[OFF_W3 + o * MH + c]) : 0;
  }
}

// Forward for samples [s0, s0 + SB) (rows >= B are zero inputs).
__device__ void forward(const float* P, const float* X, int B, int s0, MlpLds& S) {
  for (int i = threadIdx.x; i < SB * KIN; i += MT) {
    const int r = i / KIN, c = i % KIN;
    S.X[r][c] = (c < MIN && s0 + r < B) ? bf(X[(int64_t)(s0 + r) * MIN + c]) : 0;
  }
  __syncthreads();
  wg_mm(SB, MH, KIN, [&](int i, int k) { return S.X[i][k]; },
        [&](int k, int j) { return S.W1[j][k]; },
        [&](int i, int j, float v) {
          const float z = v + P[OFF_B1 + j];
          S.H1[i][j] = bf(gelu(z));
          S.G1[i][j] = gelu_d(z);
        });
  __syncthreads();
  wg_mm(SB, MH, MH, [&](int i, int k) { return S.H1[i][k]; },
        [&](int k, int j) { return S.W2[j][k]; },
        [&](int i, int j, float v) {
          const float z = v + P[OFF_B2 + j];
          S.H2[i][j] = bf(gelu(z));
          S.G2[i][j] = gelu_d(z);
        });
  __syncthreads();
  wg_mm(SB, NOUT, MH, [&](int i, int k) { return S.H2[i][k]; },
        [&](int k, int j) { return S.W3[j][k]; },
        [&](int i, int j, float v) { S.Y[i][j] = v + (j < MOUT ? P[OFF_B3 + j] : 0.f); });
  __syncthreads();
}

__global__ __launch_bounds__(MT) void mlp_forward_kernel(const float* P, const float* X, int B,
                                                         float* Yout) {
  extern __shared__ __align__(16) unsigned char raw[];
  MlpLds& S = *(MlpLds*)raw;
  stage_weights(P, S);
  const int s0 = blockIdx.x * SB;
  forward(P, X, B, s0, S);
  for (int i = threadIdx.x; i < SB * MOUT; i += MT) {
    const int r = i / MOUT, o = i % MOUT;
    if (s0 + r < B) Yout[(int64_t)(s0 + r) * MOUT + o] = S.Y[r][o];
  }
}

// Loss = sum_s sum_o wo[o] (Y_so - T_so)^2 / B.  Per work-group: its loss
// share and its weight-gradient share into grad_part + blockIdx.x * NPARAM.
__global__ __launch_bounds__(MT) void mlp_grad_kernel(const float* P, const float* X,
                                                      const float* T, const float* wo, int B,
                                                      float* grad_part, float* loss_part) {
  extern __shared__ __align__(16) unsigned char raw[];
  MlpLds& S = *(MlpLds*)raw;
  stage_weights(P, S);
  const int s0 = blockIdx.x * SB;
  forward(P, X, B, s0, S);
  float* g = grad_part + (int64_t)blockIdx.x * NPARAM;
  // dY (zero for padded samples and outputs)
  float lsum = 0.f;
  for (int i = threadIdx.x; i < SB * NOUT; i += MT) {
    const int r = i / NOUT, o = i % NOUT;
    float d = 0.f;
    if (o < MOUT && s0 + r < B) {
      const float e = S.Y[r][o] - T[(int64_t)(s0 + r) * MOUT + o];
      lsum += wo[o] * e * e;
      d = 2.f * wo[o] * e / (float)B;
    }
    S.DY[r][o] = bf(d);
  }
  lsum = block_sum(lsum, S.red);
  if (threadIdx.x == 0) loss_part[blockIdx.x] = lsum / (float)B;
  __syncthreads();
  // dW3 = dY^T H2 (16 x 128, rows >= 3 dropped); db3 = sum dY
  wg_mm(NOUT, MH, SB, [&](int i, int k) { return S.DY[k][i]; },
        [&](int k, int j) { return S.H2[k][j]; },
        [&](int i, int j, float v) { if (i < MOUT) g[OFF_W3 + i * MH + j] = v; });
  for (int o = threadIdx.x; o < MOUT; o += MT) {
    float sacc = 0.f;
    for (int r = 0; r < SB; ++r) sacc += fb(S.DY[r][o]);
    g[OFF_B3 + o] = sacc;
  }
  __syncthreads();
  // dZ2 = (dY W3) * GELU'(Z2) -> bf16 in H2
  wg_mm(SB, MH, NOUT, [&](int i, int k) { return S.DY[i][k]; },
        [&](int k, int j) { return S.W3[k][j]; },
        [&](int i, int j, float v) { S.H2[i][j] = bf(v * S.G2[i][j]); });
  __syncthreads();
  // dW2 = dZ2^T H1 ; db2 = sum dZ2
  wg_mm(MH, MH, SB, [&](int i, int k) { return S.H2[k][i]; },
        [&](int k, int j) { return S.H1[k][j]; },
        [&](int i, int j, float v) { g[OFF_W2 + i * MH + j] = v; });
  for (int j = threadIdx.x; j < MH; j += MT) {
    float sacc = 0.f;
    for (int r = 0; r < SB; ++r) sacc += fb(S.H2[r][j]);
    g[OFF_B2 + j] = sacc;
  }
  __syncthreads();
  // dZ1 = (dZ2 W2) * GELU'(Z1) -> bf16 in H1
  wg_mm(SB, MH, MH, [&](int i, int k) { return S.H2[i][k]; },
        [&](int k, int j) { return S.W2[k][j]; },
        [&](int i, int j, float v) { S.H1[i][j] = bf(v * S.G1[i][j]); });
  __syncthreads();
  // dW1 = dZ1^T X (128 x 48; columns >= 34 dropped) ; db1 = sum dZ1
  wg_mm(MH, KIN, SB, [&](int i, int k) { return S.H1[k][i]; },
        [&](int k, int j) { return S.X[k][j]; },
        [&](int i, int j, float v) { if (j < MIN) g[OFF_W1 + i * MIN + j] = v; });
  for (int j = threadIdx.x; j < MH; j += MT) {
    float sacc = 0.f;
    for (int r = 0; r < SB; ++r) sacc += fb(S.H1[r][j]);
    g[OFF_B1 + j] = sacc;
  }
}

// grad = sum over work-groups (fixed order); loss likewise.
__global__ void mlp_reduce_kernel(const float* part, int nwg, float* grad, const float* loss_part,
                                  float* loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NPARAM) {
    float s = 0.f;
    for (int w = 0; w < nwg; ++w) s += part[(int64_t)w * NPARAM + i];
    grad[i] = s;
  }
  if (i == 0 && loss) {
    float s = 0.f;
    for (int w = 0; w < nwg; ++w) s += loss_part[w];
    *loss = s;
  }
}

// Adam (bias-corrected), FP32.
__global__ void mlp_adam_kernel(float* P, float* m, float* v, const float* g, int step, float lr,
                                float b1, float b2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, (float)step));
  const float vh = vi / (1.f - powf(b2, (float)step));
  P[i] -= lr * mh / (sqrtf(vh) + eps);
}

size_t mlp_lds() { return sizeof(MlpLds); }

bool mlp_attr_done = false;
hipError_t mlp_attr() {
  if (mlp_attr_done) return hipSuccess;
  const void* fns[] = {(const void*)mlp_forward_kernel, (const void*)mlp_grad_kernel};
  for (const void* f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)mlp_lds());
    if (e != hipSuccess) return e;
  }
  mlp_attr_done = true;
  return hipSuccess;
}

}  // namespace

int mlp_param_count() { return NPARAM; }
int mlp_group_samples() { return SB; }

hipError_t mlp_forward(const float* P, const float* X, int B, float* Y, hipStream_t st) {
  if (B <= 0) return hipSuccess;
  hipError_t e = mlp_attr();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((B + SB - 1) / SB), dim3(MT), mlp_lds(), st, P, X, B,
                     Y);
  return hipGetLastError();
}

// grad (NPARAM) and loss (1) of the weighted MSE; work: nwg * (NPARAM + 1)
// floats with nwg = ceil(B / 32).
hipError_t mlp_grad(const float* P, const float* X, const float* T, const float* wo, int B,
                    float* grad, float* loss, float* work, hipStream_t st) {
  if (B <= 0) return hipErrorInvalidValue;
  hipError_t e = mlp_attr();
  if (e != hipSuccess) return e;
  const int nwg = (B + SB - 1) / SB;
  float* lp = work + (int64_t)nwg * NPARAM;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(nwg), dim3(MT), mlp_lds(), st, P, X, T, wo, B, work, lp);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, work, nwg,
                     grad, lp, loss);
  return hipGetLastError();
}

hipError_t mlp_adam(float* P, float* m, float* v, const float* g, int step, float lr, float b1,
                    float b2, float eps, hipStream_t st) {
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, P, m, v, g,
                     step, lr, b1, b2, eps);
  return hipGetLastError();
}

}  // namespace rna
