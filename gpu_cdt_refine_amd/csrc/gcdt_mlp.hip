// Learned insertion priority: a 3-layer MLP (16 -> 64 -> 64 -> 1, ReLU) with
// BF16 weights and activations feeding v_mfma_f32_16x16x16_bf16 (FP32
// accumulate), trained on the GPU by the forward/backward kernels below.
//
// Inputs are per-candidate features the cavity kernel records (minimum angle,
// circumradius / shortest-edge ratio, cavity and ring size, distance to the
// nearest cavity subsegment, pass index); the target is the candidate's
// conflict degree from the claim logs (how many other candidates claim the
// triangles of its cavity and ring).  The prediction orders candidates inside
// a priority bucket; every insertion is still decided by exact predicates.
//
// Weight layout (FP32 master copy, GCDT_MLP_NPARAM floats):
//   W1[16][64] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3
// A wave (64 lanes) processes 16 samples per step.  MFMA operand layout for
// 16x16x16 (verified by tests/test_gpu_mlp.py against a host product):
//   A[16][16]: lane l holds row l%16, k = 4*(l/16) + 0..3
//   B[16][16]: lane l holds col l%16, k = 4*(l/16) + 0..3
//   D[16][16]: lane l holds col l%16, row = 4*(l/16) + 0..3
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu_cdt_refine.h"

namespace gcdt {

constexpr int MLP_IN = 16, MLP_H = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_IN * MLP_H, OFF_W2 = OFF_B1 + MLP_H,
              OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H,
              OFF_B3 = OFF_W3 + MLP_H, MLP_NPARAM = OFF_B3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) {
  return __uint_as_float(((unsigned)(unsigned short)h) << 16);
}

__device__ __forceinline__ floatx4 mfma(bf16x4 a, bf16x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// gfx950's double-K BF16 MFMA (v_mfma_f32_16x16x32_bf16), used by the
// inference kernel's 64-wide hidden layer: two instructions per 16x16 output
// block instead of four.  Operand layout (checked by k_mfma_probe32 against a
// host product): A[16][32]: lane l holds row l%16, k = 8*(l/16) + 0..7;
// B[32][16]: lane l holds col l%16, k = 8*(l/16) + 0..7; D as for 16x16x16.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short shortx8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ floatx4 mfma32(shortx8 a, shortx8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// One wave computes the forward pass of 16 samples starting at row0.
// Outputs D-layout activations (pre-ReLU kept implicitly by value > 0).
struct Fwd {
  floatx4 h1[4], h2[4];  // post-ReLU activations, D layout, 4 column blocks
  float y[4];            // lane l: outputs of rows 4*(l/16)+j (all 16 lanes of a group hold the same)
};

__device__ __forceinline__ void mlp_forward16(const float *__restrict__ X,
                                              int n, int row0,
                                              const float *__restrict__ P,
                                              short *lds /* 16*64 */,
                                              Fwd &f) {
  int l = threadIdx.x & 63;
  int r = l % 16, kb = l / 16;
  // layer 1: A = X[row0 + r][4kb..4kb+3]
  bf16x4 a;
  for (int j = 0; j < 4; j++) {
    int row = row0 + r;
    a[j] = row < n ? f2bf(X[(size_t)row * MLP_IN + 4 * kb + j]) : (short)0;
  }
  for (int cb = 0; cb < 4; cb++) {
    bf16x4 b;
    for (int j = 0; j < 4; j++) b[j] = f2bf(P[OFF_W1 + (4 * kb + j) * MLP_H + cb * 16 + r]);
    floatx4 z = {0.f, 0.f, 0.f, 0.f};
    z = mfma(a, b, z);
    float bias = P[OFF_B1 + cb * 16 + r];
    for (int j = 0; j < 4; j++) z[j] = fmaxf(z[j] + bias, 0.f);
    f.h1[cb] = z;
  }
  // stage H1 (bf16) through LDS to re-read it in A layout
  for (int cb = 0; cb < 4; cb++)
    for (int j = 0; j < 4; j++) lds[(4 * kb + j) * MLP_H + cb * 16 + r] = f2bf(f.h1[cb][j]);
  __syncthreads();
  for (int cb = 0; cb < 4; cb++) {
    floatx4 z = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < 4; kk++) {  // K blocks of 16 over the 64 hidden units
      bf16x4 a2, b2;
      for (int j = 0; j < 4; j++) {
        a2[j] = lds[r * MLP_H + kk * 16 + 4 * kb + j];
        b2[j] = f2bf(P[OFF_W2 + (kk * 16 + 4 * kb + j) * MLP_H + cb * 16 + r]);
      }
      z = mfma(a2, b2, z);
    }
    float bias = P[OFF_B2 + cb * 16 + r];
    for (int j = 0; j < 4; j++) z[j] = fmaxf(z[j] + bias, 0.f);
    f.h2[cb] = z;
  }
  __syncthreads();
  // layer 3: y[row] = sum_c H2[row][c] W3[c] + b3, reduced over the 16 lanes
  // of a group (they hold the 16 columns of each block)
  for (int j = 0; j < 4; j++) {
    float s = 0.f;
    for (int cb = 0; cb < 4; cb++) s += f.h2[cb][j] * bf2f(f2bf(P[OFF_W3 + cb * 16 + r]));
    for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 16);
    f.y[j] = s + P[OFF_B3];
  }
}

// dn (optional): the row count is read from device memory, and the kernel
// does nothing unless *dgo is set (graph-driven refinement loop).
__global__ __launch_bounds__(64) void k_mlp_forward(const float *X, int n,
                                                    const float *P, float *y,
                                                    const int *dn, const int *dgo) {
  __shared__ short lds[16 * MLP_H];
  if (dn) n = *dgo ? *dn : 0;
  int l = threadIdx.x;
  for (int row0 = blockIdx.x * 16; row0 < n; row0 += gridDim.x * 16) {
    Fwd f;
    mlp_forward16(X, n, row0, P, lds, f);
    if (l % 16 == 0)
      for (int j = 0; j < 4; j++) {
        int row = row0 + 4 * (l / 16) + j;
        if (row < n) y[row] = f.y[j];
      }
    __syncthreads();
  }
}

// Inference: the forward pass above with the weights converted to BF16 once
// per block and kept in LDS, four waves per block, each wave taking 16 rows
// at a time; layer 2 on the 16x16x32 MFMA.  Same roundings as the training
// forward (inputs, weights and hidden activations in BF16, FP32 accumulate);
// only layer 2's accumulation order differs.
constexpr int INF_WAVES = 4;
// LDS written by some lanes of a wave and read by others: release, the
// wave's convergence point, acquire (wavefront scope).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
constexpr int W2T_STRIDE = MLP_H + 8;  // transposed W2 rows, padded against bank conflicts
constexpr int H1_STRIDE = MLP_H + 8;
__global__ __launch_bounds__(64 * INF_WAVES) void k_mlp_infer(const float *X, int n,
                                                              const float *P, float *y,
                                                              const int *dn, const int *dgo) {
  __shared__ short w1[MLP_IN * MLP_H];                 // [k][c]
  __shared__ __align__(16) short w2t[MLP_H * W2T_STRIDE];  // [c][k]
  __shared__ float b1[MLP_H], b2[MLP_H], w3[MLP_H];
  __shared__ __align__(16) short h1[INF_WAVES][16 * H1_STRIDE];  // per wave: [row][k]
  if (dn) n = *dgo ? *dn : 0;
  if (n <= 0) return;
  for (int i = threadIdx.x; i < MLP_IN * MLP_H; i += blockDim.x) w1[i] = f2bf(P[OFF_W1 + i]);
  for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) {
    const int k = i / MLP_H, c = i % MLP_H;
    w2t[c * W2T_STRIDE + k] = f2bf(P[OFF_W2 + i]);
  }
  for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) {
    b1[i] = P[OFF_B1 + i];
    b2[i] = P[OFF_B2 + i];
    w3[i] = bf2f(f2bf(P[OFF_W3 + i]));
  }
  __syncthreads();
  const float b3 = P[OFF_B3];
  const int wv = threadIdx.x / 64, l = threadIdx.x & 63, r = l % 16, kb = l / 16;
  short *H = h1[wv];
  for (int row0 = (blockIdx.x * INF_WAVES + wv) * 16; row0 < n;
       row0 += gridDim.x * INF_WAVES * 16) {
    // layer 1 (K = 16): A = X[row0 + r][4kb..4kb+3]
    bf16x4 a;
    const int row = row0 + r;
    for (int j = 0; j < 4; j++)
      a[j] = row < n ? f2bf(X[(size_t)row * MLP_IN + 4 * kb + j]) : (short)0;
    for (int cb = 0; cb < 4; cb++) {
      bf16x4 b;
      for (int j = 0; j < 4; j++) b[j] = w1[(4 * kb + j) * MLP_H + cb * 16 + r];
      floatx4 z = {0.f, 0.f, 0.f, 0.f};
      z = mfma(a, b, z);
      const float bias = b1[cb * 16 + r];
      // D layout: column cb*16 + r, rows 4kb + j
      for (int j = 0; j < 4; j++) H[(4 * kb + j) * H1_STRIDE + cb * 16 + r] = f2bf(fmaxf(z[j] + bias, 0.f));
    }
    wave_sync();  // H written by the wave's lanes, read by others
    // layer 2 (K = 64 = 2 x 32): A row r, k = kk*32 + 8kb + 0..7 (contiguous in H)
    floatx4 h2[4];
    for (int cb = 0; cb < 4; cb++) {
      floatx4 z = {0.f, 0.f, 0.f, 0.f};
      for (int kk = 0; kk < 2; kk++) {
        const shortx8 a2 = *reinterpret_cast<const shortx8 *>(&H[r * H1_STRIDE + kk * 32 + 8 * kb]);
        const shortx8 b2v =
            *reinterpret_cast<const shortx8 *>(&w2t[(cb * 16 + r) * W2T_STRIDE + kk * 32 + 8 * kb]);
        z = mfma32(a2, b2v, z);
      }
      const float bias = b2[cb * 16 + r];
      for (int j = 0; j < 4; j++) z[j] = fmaxf(z[j] + bias, 0.f);
      h2[cb] = z;
    }
    // layer 3: y[row] = sum_c H2[row][c] W3[c] + b3 over the 16 lanes of a group
    for (int j = 0; j < 4; j++) {
      float s = 0.f;
      for (int cb = 0; cb < 4; cb++) s += h2[cb][j] * w3[cb * 16 + r];
      for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 16);
      const int ro = row0 + 4 * kb + j;
      if (r == 0 && ro < n) y[ro] = s + b3;
    }
    wave_sync();  // H is rewritten by the next rows
  }
}

// Mean-squared-error gradient over n samples, accumulated into G (MLP_NPARAM
// floats, zeroed by the caller) and the loss sum into *loss.
__global__ __launch_bounds__(64) void k_mlp_grad(const float *X, const float *t,
                                                 int n, const float *P, float *G,
                                                 float *loss) {
  __shared__ short lds[16 * MLP_H];
  __shared__ float gsh[MLP_NPARAM];
  int l = threadIdx.x, r = l % 16, kb = l / 16;
  for (int k = l; k < MLP_NPARAM; k += 64) gsh[k] = 0.f;
  __syncthreads();
  float lsum = 0.f;
  for (int row0 = blockIdx.x * 16; row0 < n; row0 += gridDim.x * 16) {
    Fwd f;
    mlp_forward16(X, n, row0, P, lds, f);
    // dy for the 4 rows of this lane group
    float dy[4];
    for (int j = 0; j < 4; j++) {
      int row = row0 + 4 * kb + j;
      float e = row < n ? f.y[j] - t[row] : 0.f;
      dy[j] = 2.f * e / (float)n;
      if (r == 0) lsum += e * e;
    }
    // layer 3: dW3[c] = sum_rows H2[row][c] dy[row]; db3 = sum dy
    for (int cb = 0; cb < 4; cb++) {
      float s = 0.f;
      for (int j = 0; j < 4; j++) s += f.h2[cb][j] * dy[j];
      atomicAdd(&gsh[OFF_W3 + cb * 16 + r], s);
    }
    if (r == 0) atomicAdd(&gsh[OFF_B3], dy[0] + dy[1] + dy[2] + dy[3]);
    // dH2 = dy * W3 (.) relu'
    floatx4 dh2[4];
    for (int cb = 0; cb < 4; cb++) {
      float w3 = bf2f(f2bf(P[OFF_W3 + cb * 16 + r]));
      for (int j = 0; j < 4; j++) dh2[cb][j] = f.h2[cb][j] > 0.f ? dy[j] * w3 : 0.f;
    }
    // dW2 = H1^T dH2: A = H1^T block (D layout of H1), B = dH2 (D layout)
    for (int mb = 0; mb < 4; mb++) {
      bf16x4 a;
      for (int j = 0; j < 4; j++) a[j] = f2bf(f.h1[mb][j]);
      for (int nb = 0; nb < 4; nb++) {
        bf16x4 b;
        for (int j = 0; j < 4; j++) b[j] = f2bf(dh2[nb][j]);
        floatx4 z = {0.f, 0.f, 0.f, 0.f};
        z = mfma(a, b, z);  // rows: hidden1 unit mb*16 + 4kb + j, col: hidden2 nb*16 + r
        for (int j = 0; j < 4; j++)
          atomicAdd(&gsh[OFF_W2 + (mb * 16 + 4 * kb + j) * MLP_H + nb * 16 + r], z[j]);
      }
      // db2: column sums of dH2
    }
    for (int cb = 0; cb < 4; cb++) {
      float s = dh2[cb][0] + dh2[cb][1] + dh2[cb][2] + dh2[cb][3];
      atomicAdd(&gsh[OFF_B2 + cb * 16 + r], s);
    }
    // dH1 = dH2 W2^T (.) relu'(H1): stage dH2 (bf16) through LDS for A layout
    for (int cb = 0; cb < 4; cb++)
      for (int j = 0; j < 4; j++) lds[(4 * kb + j) * MLP_H + cb * 16 + r] = f2bf(dh2[cb][j]);
    __syncthreads();
    floatx4 dh1[4];
    for (int cb = 0; cb < 4; cb++) {  // output: hidden1 block cb
      floatx4 z = {0.f, 0.f, 0.f, 0.f};
      for (int kk = 0; kk < 4; kk++) {  // K: hidden2 units
        bf16x4 a2, b2;
        for (int j = 0; j < 4; j++) {
          a2[j] = lds[r * MLP_H + kk * 16 + 4 * kb + j];
          // B[k][n] = W2^T[k][n] = W2[n][k], n = hidden1 unit cb*16 + r
          b2[j] = f2bf(P[OFF_W2 + (cb * 16 + r) * MLP_H + kk * 16 + 4 * kb + j]);
        }
        z = mfma(a2, b2, z);
      }
      for (int j = 0; j < 4; j++) z[j] = f.h1[cb][j] > 0.f ? z[j] : 0.f;
      dh1[cb] = z;
    }
    __syncthreads();
    // dW1 = X^T dH1: A = X^T (lane l: feature l%16, samples 4kb+j), B = dH1 (D layout)
    bf16x4 ax;
    for (int j = 0; j < 4; j++) {
      int row = row0 + 4 * kb + j;
      ax[j] = row < n ? f2bf(X[(size_t)row * MLP_IN + r]) : (short)0;
    }
    for (int nb = 0; nb < 4; nb++) {
      bf16x4 b;
      for (int j = 0; j < 4; j++) b[j] = f2bf(dh1[nb][j]);
      floatx4 z = {0.f, 0.f, 0.f, 0.f};
      z = mfma(ax, b, z);  // rows: feature 4kb+j, col: hidden1 nb*16 + r
      for (int j = 0; j < 4; j++)
        atomicAdd(&gsh[OFF_W1 + (4 * kb + j) * MLP_H + nb * 16 + r], z[j]);
      float s = dh1[nb][0] + dh1[nb][1] + dh1[nb][2] + dh1[nb][3];
      atomicAdd(&gsh[OFF_B1 + nb * 16 + r], s);
    }
    __syncthreads();
  }
  __syncthreads();
  for (int k = l; k < MLP_NPARAM; k += 64)
    if (gsh[k] != 0.f) atomicAdd(&G[k], gsh[k]);
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (l == 0) atomicAdd(loss, lsum);
}

// Adam step on the FP32 master weights.
__global__ void k_adam(float *P, const float *G, float *m1, float *m2, int np,
                       float lr, float b1, float b2, float eps, int step) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  float g = G[k];
  float a = b1 * m1[k] + (1.f - b1) * g;
  float v = b2 * m2[k] + (1.f - b2) * g * g;
  m1[k] = a;
  m2[k] = v;
  float ah = a / (1.f - powf(b1, (float)step));
  float vh = v / (1.f - powf(b2, (float)step));
  P[k] -= lr * ah / (sqrtf(vh) + eps);
}

// Device MFMA layout probe: D = A(16x16) B(16x16) for host verification.
__global__ void k_mfma_probe(const float *A, const float *B, float *D) {
  int l = threadIdx.x, r = l % 16, kb = l / 16;
  bf16x4 a, b;
  for (int j = 0; j < 4; j++) {
    a[j] = f2bf(A[r * 16 + 4 * kb + j]);
    b[j] = f2bf(B[(4 * kb + j) * 16 + r]);
  }
  floatx4 z = {0.f, 0.f, 0.f, 0.f};
  z = mfma(a, b, z);
  for (int j = 0; j < 4; j++) D[(4 * kb + j) * 16 + r] = z[j];
}

// Device layout probe of the 16x16x32 MFMA: D = A(16x32) B(32x16).
__global__ void k_mfma_probe32(const float *A, const float *B, float *D) {
  int l = threadIdx.x, r = l % 16, kb = l / 16;
  shortx8 a, b;
  for (int j = 0; j < 8; j++) {
    a[j] = f2bf(A[r * 32 + 8 * kb + j]);
    b[j] = f2bf(B[(8 * kb + j) * 16 + r]);
  }
  floatx4 z = {0.f, 0.f, 0.f, 0.f};
  z = mfma32(a, b, z);
  for (int j = 0; j < 4; j++) D[(4 * kb + j) * 16 + r] = z[j];
}

static void hc(hipError_t e, const char *w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

// Inference inside the refinement: scores for n candidates on `stream`.
// Blocks of the inference kernel for n rows: enough to fill the chip, each
// block converting the weights once.
static int infer_blocks(int n) {
  return std::max(1, std::min((n + 64 * 8 - 1) / (64 * 8), 1024));
}

void mlp_forward_device(const float *dX, int n, const float *dP, float *dy,
                        hipStream_t stream) {
  if (n > 0) k_mlp_infer<<<infer_blocks(n), 64 * INF_WAVES, 0, stream>>>(dX, n, dP, dy, nullptr, nullptr);
  hc(hipGetLastError(), "mlp forward launch");
}

void mlp_forward_device_n(const float *dX, const int *dn, const int *dgo,
                          const float *dP, float *dy, int blocks,
                          hipStream_t stream) {
  k_mlp_infer<<<std::min(blocks, 1024), 64 * INF_WAVES, 0, stream>>>(dX, 0, dP, dy, dn, dgo);
  hc(hipGetLastError(), "mlp forward launch");
}

}  // namespace gcdt

using namespace gcdt;

static thread_local std::string g_mlp_err;

extern "C" {

int gcdt_mlp_nparam(void) { return MLP_NPARAM; }

// y[n] = MLP(X[n][16]) on the device.
int gcdt_mlp_forward(const float *X, int n, const float *params, float *y) {
  try {
    float *dX, *dP, *dy;
    hc(hipMalloc(&dX, sizeof(float) * (size_t)std::max(n, 1) * MLP_IN), "malloc");
    hc(hipMalloc(&dP, sizeof(float) * MLP_NPARAM), "malloc");
    hc(hipMalloc(&dy, sizeof(float) * (size_t)std::max(n, 1)), "malloc");
    hc(hipMemcpy(dX, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice), "copy");
    hc(hipMemcpy(dP, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice), "copy");
    if (n > 0) k_mlp_infer<<<infer_blocks(n), 64 * INF_WAVES>>>(dX, n, dP, dy, nullptr, nullptr);
    hc(hipGetLastError(), "launch");
    hc(hipDeviceSynchronize(), "sync");
    hc(hipMemcpy(y, dy, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost), "copy");
    (void)hipFree(dX); (void)hipFree(dP); (void)hipFree(dy);
    return 0;
  } catch (const std::exception &e) {
    g_mlp_err = e.what();
    return -2;
  }
}

// Mean-squared-error gradient (grad[MLP_NPARAM]) and mean loss on the device.
int gcdt_mlp_grad(const float *X, const float *t, int n, const float *params,
                  float *grad, float *loss) {
  try {
    float *dX, *dt, *dP, *dG, *dl;
    hc(hipMalloc(&dX, sizeof(float) * (size_t)std::max(n, 1) * MLP_IN), "malloc");
    hc(hipMalloc(&dt, sizeof(float) * (size_t)std::max(n, 1)), "malloc");
    hc(hipMalloc(&dP, sizeof(float) * MLP_NPARAM), "malloc");
    hc(hipMalloc(&dG, sizeof(float) * MLP_NPARAM), "malloc");
    hc(hipMalloc(&dl, sizeof(float)), "malloc");
    hc(hipMemcpy(dX, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice), "copy");
    hc(hipMemcpy(dt, t, sizeof(float) * (size_t)n, hipMemcpyHostToDevice), "copy");
    hc(hipMemcpy(dP, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice), "copy");
    hc(hipMemset(dG, 0, sizeof(float) * MLP_NPARAM), "memset");
    hc(hipMemset(dl, 0, sizeof(float)), "memset");
    int blocks = std::min((n + 15) / 16, 1024);
    if (blocks > 0) k_mlp_grad<<<blocks, 64>>>(dX, dt, n, dP, dG, dl);
    hc(hipGetLastError(), "launch");
    hc(hipDeviceSynchronize(), "sync");
    hc(hipMemcpy(grad, dG, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost), "copy");
    float ls;
    hc(hipMemcpy(&ls, dl, sizeof(float), hipMemcpyDeviceToHost), "copy");
    *loss = ls / (float)std::max(n, 1);
    (void)hipFree(dX); (void)hipFree(dt); (void)hipFree(dP); (void)hipFree(dG); (void)hipFree(dl);
    return 0;
  } catch (const std::exception &e) {
    g_mlp_err = e.what();
    return -2;
  }
}

// Train params (in/out) with Adam on the device: mini-batches of `batch`
// samples drawn in a seeded shuffled order, `epochs` passes.  losses[epoch]
// receives the mean training loss of each epoch (may be NULL).
int gcdt_mlp_train(const float *X, const float *t, int n, int epochs, int batch,
                   float lr, unsigned long long seed, float *params,
                   float *losses) {
  try {
    if (n <= 0 || batch <= 0) throw std::runtime_error("empty training set");
    // shuffle once per epoch on the host (index permutation), gather on host
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    unsigned long long s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {
      s ^= s << 13;
      s ^= s >> 7;
      s ^= s << 17;
      return s;
    };
    float *dX, *dt, *dP, *dG, *dm1, *dm2, *dl;
    hc(hipMalloc(&dX, sizeof(float) * (size_t)n * MLP_IN), "malloc");
    hc(hipMalloc(&dt, sizeof(float) * (size_t)n), "malloc");
    hc(hipMalloc(&dP, sizeof(float) * MLP_NPARAM), "malloc");
    hc(hipMalloc(&dG, sizeof(float) * MLP_NPARAM), "malloc");
    hc(hipMalloc(&dm1, sizeof(float) * MLP_NPARAM), "malloc");
    hc(hipMalloc(&dm2, sizeof(float) * MLP_NPARAM), "malloc");
    hc(hipMalloc(&dl, sizeof(float)), "malloc");
    hc(hipMemcpy(dP, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice), "copy");
    hc(hipMemset(dm1, 0, sizeof(float) * MLP_NPARAM), "memset");
    hc(hipMemset(dm2, 0, sizeof(float) * MLP_NPARAM), "memset");
    std::vector<float> hX((size_t)n * MLP_IN), ht(n);
    int step = 0;
    for (int ep = 0; ep < epochs; ep++) {
      for (int i = n - 1; i > 0; i--) std::swap(perm[i], perm[rnd() % (unsigned long long)(i + 1)]);
      for (int i = 0; i < n; i++) {
        std::memcpy(&hX[(size_t)i * MLP_IN], &X[(size_t)perm[i] * MLP_IN], sizeof(float) * MLP_IN);
        ht[i] = t[perm[i]];
      }
      hc(hipMemcpy(dX, hX.data(), sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice), "copy");
      hc(hipMemcpy(dt, ht.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice), "copy");
      hc(hipMemset(dl, 0, sizeof(float)), "memset");
      double eloss = 0.0;
      for (int b0 = 0; b0 < n; b0 += batch) {
        int nb = std::min(batch, n - b0);
        hc(hipMemset(dG, 0, sizeof(float) * MLP_NPARAM), "memset");
        k_mlp_grad<<<std::min((nb + 15) / 16, 256), 64>>>(dX + (size_t)b0 * MLP_IN, dt + b0, nb, dP, dG, dl);
        step++;
        k_adam<<<(MLP_NPARAM + 255) / 256, 256>>>(dP, dG, dm1, dm2, MLP_NPARAM, lr, 0.9f, 0.999f, 1e-8f, step);
      }
      hc(hipGetLastError(), "launch");
      float ls;
      hc(hipMemcpy(&ls, dl, sizeof(float), hipMemcpyDeviceToHost), "copy");
      eloss = ls;
      // k_mlp_grad accumulates e^2 / n_batch-weighted sums: report per-sample mean
      if (losses) losses[ep] = (float)(eloss / (double)n);
    }
    hc(hipMemcpy(params, dP, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost), "copy");
    (void)hipFree(dX); (void)hipFree(dt); (void)hipFree(dP); (void)hipFree(dG);
    (void)hipFree(dm1); (void)hipFree(dm2); (void)hipFree(dl);
    return 0;
  } catch (const std::exception &e) {
    g_mlp_err = e.what();
    return -2;
  }
}

// D = A B for 16x16 FP32 matrices rounded to BF16 (MFMA layout probe).
// D(16x16) = A(16x32) B(32x16) through the 16x16x32 MFMA (layout check).
int gcdt_mfma_probe32(const float *A, const float *B, float *D) {
  try {
    float *dA, *dB, *dD;
    hc(hipMalloc(&dA, 2048), "malloc");
    hc(hipMalloc(&dB, 2048), "malloc");
    hc(hipMalloc(&dD, 1024), "malloc");
    hc(hipMemcpy(dA, A, 2048, hipMemcpyHostToDevice), "copy");
    hc(hipMemcpy(dB, B, 2048, hipMemcpyHostToDevice), "copy");
    k_mfma_probe32<<<1, 64>>>(dA, dB, dD);
    hc(hipGetLastError(), "launch");
    hc(hipDeviceSynchronize(), "sync");
    hc(hipMemcpy(D, dD, 1024, hipMemcpyDeviceToHost), "copy");
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dD);
    return 0;
  } catch (const std::exception &e) {
    g_mlp_err = e.what();
    return -2;
  }
}

int gcdt_mfma_probe(const float *A, const float *B, float *D) {
  try {
    float *dA, *dB, *dD;
    hc(hipMalloc(&dA, 1024), "malloc");
    hc(hipMalloc(&dB, 1024), "malloc");
    hc(hipMalloc(&dD, 1024), "malloc");
    hc(hipMemcpy(dA, A, 1024, hipMemcpyHostToDevice), "copy");
    hc(hipMemcpy(dB, B, 1024, hipMemcpyHostToDevice), "copy");
    k_mfma_probe<<<1, 64>>>(dA, dB, dD);
    hc(hipGetLastError(), "launch");
    hc(hipDeviceSynchronize(), "sync");
    hc(hipMemcpy(D, dD, 1024, hipMemcpyDeviceToHost), "copy");
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dD);
    return 0;
  } catch (const std::exception &e) {
    g_mlp_err = e.what();
    return -2;
  }
}

const char *gcdt_mlp_last_error(void) { return g_mlp_err.c_str(); }

}  // extern "C"
