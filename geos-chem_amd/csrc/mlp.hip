// Learned initial step size and stiffness class of a cell: a 40-input MLP
// with three hidden layers of 128 GELU units and 5 outputs (log h0, four
// class logits).  No reference counterpart: the reference starts each cell's
// Rosenbrock integration from RCNTRL(3) = KPPHvalue, the previous step's
// last step size (GeosCore/fullchem_mod.F90:1016-1020), or 1e-5 s cold.  The
// prediction only sets Hstart; the adaptive controller still accepts or
// rejects every step, so the tolerance is met as without it.
//
// Inference: bf16 MFMA (v_mfma_f32_16x16x16_bf16), fp32 accumulation, the
// activations kept in registers between layers by computing each layer
// transposed (units x cells), so one layer's accumulator tile is the next
// layer's B operand.  Training: fp32 kernels (LDS-tiled GEMM, GELU,
// loss, column sums, Adam).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "cell_fields.h"
#include "gckpp_internal.h"
#include "mlp.h"

namespace gckpp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (finite inputs; NaN stays NaN).
__device__ __forceinline__ unsigned short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Feature f of a cell, normalised.
__device__ __forceinline__ float mlp_feature(int f, const MlpDesc &d, const double *met,
                                             const double *photol, const double *conc,
                                             const double *hprev, int64_t n, int64_t cell) {
  double v;
  const double hp = hprev ? hprev[cell] : 0.0;
  if (f == 0) v = met[(int64_t)MET_TEMP * n + cell];
  else if (f == 1) v = log10(fmax(met[(int64_t)MET_PRESS * n + cell], 1e-30));
  else if (f == 2) v = log10(fmax(met[(int64_t)MET_NUMDEN * n + cell], 1e-30));
  else if (f == 3) v = met[(int64_t)MET_SUNCOS * n + cell];
  else if (f < 4 + MLP_NJ) v = log10(fmax(photol[(int64_t)d.jslot[f - 4] * n + cell], 0.0) + 1e-20);
  else if (f < 4 + MLP_NJ + MLP_NS)
    v = log10(fmax(conc[(int64_t)d.spc[f - 4 - MLP_NJ] * n + cell], 0.0) + 1.0);
  else if (f == MLP_NIN - 2) v = hp > 0.0 ? log(hp) : 0.0;
  else if (f == MLP_NIN - 1) v = hp > 0.0 ? 1.0 : 0.0;
  else return 0.0f;
  return (float)((v - d.mean[f]) * d.rstd[f]);
}

// X[cell][MLP_F] (fp32, padded features zero).
__global__ __launch_bounds__(256) void k_mlp_features(int64_t n, MlpDesc d, const double *met,
                                                      const double *photol, const double *conc,
                                                      const double *hprev, float *X) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * MLP_F) return;
  int64_t cell = i / MLP_F;
  int f = (int)(i % MLP_F);
  X[i] = mlp_feature(f, d, met, photol, conc, hprev, n, cell);
}

// Packs W[in][out] (fp32, row-major) of one layer into MFMA A-operand tiles of
// W^T: tile (m, c) holds W^T[16m .. 16m+15][16c .. 16c+15]; lane l of the
// tile holds 4 bf16 at W^T[16m + l%16][16c + 4(l/16) + j], j = 0..3.
__global__ void k_mlp_pack(const float *W, int nin, int nout, unsigned short *pk) {
  int kc = nin / 16, mt = nout / 16;
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // (m, c, lane)
  if (i >= mt * kc * 64) return;
  int l = i % 64, c = (i / 64) % kc, m = i / (64 * kc);
  for (int j = 0; j < 4; ++j) {
    int in = 16 * c + 4 * (l / 16) + j, out = 16 * m + l % 16;
    pk[(int64_t)i * 4 + j] = to_bf16(W[(int64_t)in * nout + out]);
  }
}

// One wave per 16 cells.  Lane l works on cell 16t + l%16 and holds rows
// 4(l/16)..4(l/16)+3 of each 16-row tile.
__global__ __launch_bounds__(256) void k_mlp_predict(int64_t n, MlpDesc d, const double *met,
                                                     const double *photol, const double *conc,
                                                     const double *hprev, const bf16x4 *pk,
                                                     const float *params,
                                                     double safety, double hmin, double hmax,
                                                     double *h0, int32_t *cls) {
  const int l = threadIdx.x & 63;
  const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int g = l >> 4;
  const int64_t cell = tile * 16 + (l & 15);
  const bool valid = cell < n;
  const int64_t cc = valid ? cell : n - 1;  // every lane runs the MFMAs

  bf16x4 b[8];
  f32x4 acc[8];
  // layer 1 input: features 16c + 4g + j
#pragma unroll
  for (int c = 0; c < MLP_F / 16; ++c)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      b[c][j] = (short)to_bf16(mlp_feature(16 * c + 4 * g + j, d, met, photol, conc, hprev, n, cc));
  const bf16x4 *pk1 = pk;
  const bf16x4 *pk2 = pk1 + (MLP_H / 16) * (MLP_F / 16) * 64;
  const bf16x4 *pk3 = pk2 + (MLP_H / 16) * (MLP_H / 16) * 64;
  const bf16x4 *pk4 = pk3 + (MLP_H / 16) * (MLP_H / 16) * 64;
  const float *b1 = params + MLP_OFF_B1, *b2 = params + MLP_OFF_B2, *b3 = params + MLP_OFF_B3,
              *b4 = params + MLP_OFF_B4;
  auto layer = [&](const bf16x4 *w, int kc, const float *bias) {
#pragma unroll
    for (int m = 0; m < MLP_H / 16; ++m) {
      f32x4 a = {0.f, 0.f, 0.f, 0.f};
      for (int c = 0; c < kc; ++c)
        a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w[(m * kc + c) * 64 + l], b[c], a, 0, 0, 0);
      acc[m] = a;
    }
#pragma unroll
    for (int m = 0; m < MLP_H / 16; ++m)
#pragma unroll
      for (int j = 0; j < 4; ++j) b[m][j] = (short)to_bf16(gelu(acc[m][j] + bias[16 * m + 4 * g + j]));
  };
  layer(pk1, MLP_F / 16, b1);
  layer(pk2, MLP_H / 16, b2);
  layer(pk3, MLP_H / 16, b3);
  f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < MLP_H / 16; ++c)
    o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(pk4[c * 64 + l], b[c], o, 0, 0, 0);
  // rows 0..3 in lanes 0-15 (log h0, logits 0..2), row 4 in lanes 16-31 (logit 3)
  float logit3 = __shfl(o[0], (l + 16) & 63) + b4[4];
  if (g == 0 && valid) {
    float lh = o[0] + b4[0];
    float z[4] = {o[1] + b4[1], o[2] + b4[2], o[3] + b4[3], logit3};
    int k = 0;
    for (int q = 1; q < 4; ++q)
      if (z[q] > z[k]) k = q;
    double h = exp((double)d.out_mu + (double)d.out_sigma * (double)lh) * safety;
    h = h < hmin ? hmin : (h > hmax ? hmax : h);
    if (h0) h0[cell] = h;
    if (cls) cls[cell] = k;
  }
}

// ---------------------------------------------------------------------------
// fp32 training kernels.
// C[M][N] = op(A) op(B) + (bias ? bias[col] : 0) + beta*C, row-major with
// leading dimensions; op(A)[i][k] = transA ? A[k*lda + i] : A[i*lda + k].
// 64x64 tiles, 256 threads (4x4 outputs each), K staged through LDS in 16s.
__global__ __launch_bounds__(256) void k_gemm_f32(int M, int N, int K, const float *A, int lda,
                                                  int transA, const float *B, int ldb, int transB,
                                                  float *C, int ldc, float beta,
                                                  const float *bias) {
  __shared__ float As[16][65], Bs[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      int kk = e / 64, ii = e % 64;
      int gi = row0 + ii, gk = k0 + kk;
      float va = 0.f;
      if (gi < M && gk < K) va = transA ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk];
      As[kk][ii] = va;
      int gj = col0 + ii;
      float vb = 0.f;
      if (gj < N && gk < K) vb = transB ? B[(int64_t)gj * ldb + gk] : B[(int64_t)gk * ldb + gj];
      Bs[kk][ii] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], bb[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ty * 4 + r];
#pragma unroll
      for (int s = 0; s < 4; ++s) bb[s] = Bs[kk][tx * 4 + s];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int s = 0; s < 4; ++s) acc[r][s] = fmaf(a[r], bb[s], acc[r][s]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      int i = row0 + ty * 4 + r, j = col0 + tx * 4 + s;
      if (i < M && j < N) {
        float v = acc[r][s];
        if (bias) v += bias[j];
        if (beta != 0.f) v += beta * C[(int64_t)i * ldc + j];
        C[(int64_t)i * ldc + j] = v;
      }
    }
}

__global__ void k_gelu_fwd(int64_t n, const float *z, float *a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = gelu(z[i]);
}
// dz = da * gelu'(z), in place on da
__global__ void k_gelu_bwd(int64_t n, const float *z, float *da) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) da[i] *= gelu_grad(z[i]);
}

// Loss and its gradient for a batch of nb rows of O[nb][16]:
//   L = wh * mean (O0 - yh)^2 + wc * mean CE(softmax(O1..O4), ycls)
// dO[nb][16]; lossp[block] = this block's part of L.
__global__ __launch_bounds__(256) void k_mlp_loss(int64_t nb, const float *O, const float *yh,
                                                  const int32_t *ycls, float wh, float wc,
                                                  float *dO, float *lossp) {
  __shared__ float red[256];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float L = 0.f;
  if (i < nb) {
    const float *o = O + i * MLP_O;
    float *d = dO + i * MLP_O;
    float e = o[0] - yh[i];
    L += wh * e * e / (float)nb;
    d[0] = 2.f * wh * e / (float)nb;
    float mx = o[1];
    for (int k = 2; k <= 4; ++k) mx = fmaxf(mx, o[k]);
    float p[4], s = 0.f;
    for (int k = 0; k < 4; ++k) {
      p[k] = expf(o[1 + k] - mx);
      s += p[k];
    }
    int y = ycls[i];
    L += wc * (logf(s) + mx - o[1 + y]) / (float)nb;
    for (int k = 0; k < 4; ++k) d[1 + k] = wc * (p[k] / s - (k == y ? 1.f : 0.f)) / (float)nb;
    for (int k = 5; k < MLP_O; ++k) d[k] = 0.f;
  }
  red[threadIdx.x] = L;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

// out[j] = sum_i X[i][j] over nrow rows (one block per column, fixed order).
__global__ __launch_bounds__(256) void k_colsum(int64_t nrow, int ncol, const float *X, float *out) {
  __shared__ float red[256];
  int j = blockIdx.x;
  float s = 0.f;
  for (int64_t i = threadIdx.x; i < nrow; i += 256) s += X[i * ncol + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

__global__ void k_adam(int n, float *p, const float *gr, float *m, float *v, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = gr[i];
  float mi = b1 * m[i] + (1.f - b1) * g;
  float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// Host-side launchers (called by the C ABI in capi.hip).
static void gemm(hipStream_t s, int M, int N, int K, const float *A, int lda, int tA, const float *B,
                 int ldb, int tB, float *C, int ldc, float beta, const float *bias) {
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL(k_gemm_f32, grid, dim3(256), 0, s, M, N, K, A, lda, tA, B, ldb, tB, C, ldc, beta,
                     bias);
}
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void mlp_pack(const float *params, unsigned short *pk, hipStream_t s) {
  const int layers[4][3] = {{MLP_OFF_W1, MLP_F, MLP_H}, {MLP_OFF_W2, MLP_H, MLP_H},
                            {MLP_OFF_W3, MLP_H, MLP_H}, {MLP_OFF_W4, MLP_H, MLP_O}};
  int64_t off = 0;
  for (auto &L : layers) {
    int tiles = (L[2] / 16) * (L[1] / 16) * 64;
    hipLaunchKernelGGL(k_mlp_pack, dim3(nblk(tiles, 256)), dim3(256), 0, s, params + L[0], L[1], L[2],
                       pk + off * 4);
    off += tiles;
  }
}

void mlp_features(const MlpDesc &d, int64_t n, const double *met, const double *photol,
                  const double *conc, const double *hprev, float *X, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_features, dim3(nblk(n * MLP_F, 256)), dim3(256), 0, s, n, d, met, photol,
                     conc, hprev, X);
}

void mlp_predict(const MlpDesc &d, int64_t n, const double *met, const double *photol,
                 const double *conc, const double *hprev, const unsigned short *pk,
                 const float *params, double safety, double hmin, double hmax, double *h0,
                 int32_t *cls, hipStream_t s) {
  int64_t tiles = (n + 15) / 16;
  hipLaunchKernelGGL(k_mlp_predict, dim3(nblk(tiles, 4)), dim3(256), 0, s, n, d, met, photol, conc,
                     hprev, (const bf16x4 *)pk, params, safety, hmin, hmax, h0, cls);
}

// fp32 forward of nb rows; act holds Z1,A1,Z2,A2,Z3,A3 ([nb][128] each) and O [nb][16].
void mlp_forward_f32(const float *P, int64_t nb, const float *X, float *act, hipStream_t s) {
  int n = (int)nb;
  float *Z1 = act, *A1 = Z1 + nb * MLP_H, *Z2 = A1 + nb * MLP_H, *A2 = Z2 + nb * MLP_H;
  float *Z3 = A2 + nb * MLP_H, *A3 = Z3 + nb * MLP_H, *O = A3 + nb * MLP_H;
  gemm(s, n, MLP_H, MLP_F, X, MLP_F, 0, P + MLP_OFF_W1, MLP_H, 0, Z1, MLP_H, 0.f, P + MLP_OFF_B1);
  hipLaunchKernelGGL(k_gelu_fwd, dim3(nblk(nb * MLP_H, 256)), dim3(256), 0, s, nb * MLP_H, Z1, A1);
  gemm(s, n, MLP_H, MLP_H, A1, MLP_H, 0, P + MLP_OFF_W2, MLP_H, 0, Z2, MLP_H, 0.f, P + MLP_OFF_B2);
  hipLaunchKernelGGL(k_gelu_fwd, dim3(nblk(nb * MLP_H, 256)), dim3(256), 0, s, nb * MLP_H, Z2, A2);
  gemm(s, n, MLP_H, MLP_H, A2, MLP_H, 0, P + MLP_OFF_W3, MLP_H, 0, Z3, MLP_H, 0.f, P + MLP_OFF_B3);
  hipLaunchKernelGGL(k_gelu_fwd, dim3(nblk(nb * MLP_H, 256)), dim3(256), 0, s, nb * MLP_H, Z3, A3);
  gemm(s, n, MLP_O, MLP_H, A3, MLP_H, 0, P + MLP_OFF_W4, MLP_O, 0, O, MLP_O, 0.f, P + MLP_OFF_B4);
}

// Gradient of the loss w.r.t. every parameter into G[MLP_NPARAM]; lossp gets
// per-block loss parts (nblk(nb,256) floats); scratch dO [nb][16], dA [nb][128].
void mlp_grad_f32(const float *P, int64_t nb, const float *X, const float *yh, const int32_t *ycls,
                  float wh, float wc, float *act, float *dO, float *dA, float *G, float *lossp,
                  hipStream_t s) {
  int n = (int)nb;
  mlp_forward_f32(P, nb, X, act, s);
  float *Z1 = act, *A1 = Z1 + nb * MLP_H, *Z2 = A1 + nb * MLP_H, *A2 = Z2 + nb * MLP_H;
  float *Z3 = A2 + nb * MLP_H, *A3 = Z3 + nb * MLP_H, *O = A3 + nb * MLP_H;
  hipLaunchKernelGGL(k_mlp_loss, dim3(nblk(nb, 256)), dim3(256), 0, s, nb, O, yh, ycls, wh, wc, dO,
                     lossp);
  // output layer
  gemm(s, MLP_H, MLP_O, n, A3, MLP_H, 1, dO, MLP_O, 0, G + MLP_OFF_W4, MLP_O, 0.f, nullptr);
  hipLaunchKernelGGL(k_colsum, dim3(MLP_O), dim3(256), 0, s, nb, MLP_O, dO, G + MLP_OFF_B4);
  gemm(s, n, MLP_H, MLP_O, dO, MLP_O, 0, P + MLP_OFF_W4, MLP_O, 1, dA, MLP_H, 0.f, nullptr);
  // hidden layers 3, 2, 1: dZ = dA * gelu'(Z); dW = A_prev^T dZ; db = colsum dZ; dA_prev = dZ W^T
  struct { float *Z; const float *Aprev; int nin; int offW, offB; } L[3] = {
      {Z3, A2, MLP_H, MLP_OFF_W3, MLP_OFF_B3},
      {Z2, A1, MLP_H, MLP_OFF_W2, MLP_OFF_B2},
      {Z1, X, MLP_F, MLP_OFF_W1, MLP_OFF_B1}};
  float *dcur = dA;
  for (int k = 0; k < 3; ++k) {
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(nb * MLP_H, 256)), dim3(256), 0, s, nb * MLP_H, L[k].Z,
                       dcur);
    gemm(s, L[k].nin, MLP_H, n, L[k].Aprev, L[k].nin, 1, dcur, MLP_H, 0, G + L[k].offW, MLP_H, 0.f,
         nullptr);
    hipLaunchKernelGGL(k_colsum, dim3(MLP_H), dim3(256), 0, s, nb, MLP_H, dcur, G + L[k].offB);
    if (k < 2) {
      // dA_prev = dZ W^T into Z's buffer (Z is not needed any more)
      gemm(s, n, MLP_H, MLP_H, dcur, MLP_H, 0, P + L[k].offW, MLP_H, 1, L[k].Z, MLP_H, 0.f, nullptr);
      dcur = L[k].Z;
    }
  }
}

void mlp_sum(const float *x, int64_t n, float *out, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, s, n, 1, x, out);
}

void mlp_adam(float *P, const float *G, float *m, float *v, float lr, int step, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(MLP_NPARAM, 256)), dim3(256), 0, s, MLP_NPARAM, P, G, m, v, lr,
                     b1, b2, eps, c1, c2);
}

}  // namespace gckpp

namespace gckpp {

// ---------------------------------------------------------------------------
// Stable counting sort of cells by class, costliest class (3) first: the
// queue order of the persistent solver.  hist[b*4 + k] = cells of class k in
// block b; one workgroup turns the counts into start offsets; each block
// then writes its cells in order.
__global__ __launch_bounds__(256) void k_class_hist(int64_t n, const int32_t *cls, int32_t *hist) {
  __shared__ int32_t cnt[4];
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) atomicAdd(&cnt[3 - min(max(cls[i], 0), 3)], 1);
  __syncthreads();
  if (threadIdx.x < 4) hist[(int64_t)blockIdx.x * 4 + threadIdx.x] = cnt[threadIdx.x];
}

// In place: hist[b*4+k] <- start of block b's class-k cells in the output.
// Thread t owns the contiguous blocks [t*per, (t+1)*per): key by key, the
// start of its first block is the totals of all earlier keys plus the
// totals of key k over the blocks of threads before it.
__global__ __launch_bounds__(256) void k_class_scan(int64_t nb, int32_t *hist) {
  __shared__ int32_t tot[4][256];
  __shared__ int32_t base[4];
  const int64_t per = (nb + 255) / 256;
  const int64_t b0 = (int64_t)threadIdx.x * per;
  const int64_t b1 = b0 + per < nb ? b0 + per : nb;
  for (int k = 0; k < 4; ++k) {
    int32_t s = 0;
    for (int64_t b = b0; b < b1; ++b) s += hist[b * 4 + k];
    tot[k][threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t run = 0;
    for (int k = 0; k < 4; ++k) {
      base[k] = run;
      for (int t = 0; t < 256; ++t) run += tot[k][t];
    }
  }
  __syncthreads();
  for (int k = 0; k < 4; ++k) {
    int32_t run = base[k];
    for (int t = 0; t < (int)threadIdx.x; ++t) run += tot[k][t];
    for (int64_t b = b0; b < b1; ++b) {
      int32_t c = hist[b * 4 + k];
      hist[b * 4 + k] = run;
      run += c;
    }
  }
}

__global__ __launch_bounds__(256) void k_class_scatter(int64_t n, const int32_t *cls,
                                                       const int32_t *start, int32_t *order) {
  __shared__ int32_t key[256];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int k = i < n ? 3 - min(max(cls[i], 0), 3) : 4;
  key[threadIdx.x] = k;
  __syncthreads();
  if (i < n) {
    int32_t r = 0;  // rank among the earlier cells of this block with the same key
    for (int t = 0; t < (int)threadIdx.x; ++t) r += key[t] == k;
    order[start[(int64_t)blockIdx.x * 4 + k] + r] = (int32_t)i;
  }
}

void order_by_class(int64_t n, const int32_t *cls, int32_t *order, int32_t *scratch, hipStream_t s) {
  int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_class_hist, dim3((unsigned)nb), dim3(256), 0, s, n, cls, scratch);
  hipLaunchKernelGGL(k_class_scan, dim3(1), dim3(256), 0, s, nb, scratch);
  hipLaunchKernelGGL(k_class_scatter, dim3((unsigned)nb), dim3(256), 0, s, n, cls, scratch, order);
}

// ---------------------------------------------------------------------------
// Stable counting sort of cells by predicted initial step size, smallest
// (the stiffest, costliest cells) first: key = one of H0_BINS bins of
// log10(h0) over [1e-5 s, 1200 s] (a quarter decade each); h0 <= 0 or not
// finite (no prediction) sorts last.  Same three passes as the class sort,
// over the keys stored in the scratch.
constexpr int H0_BINS = 32;
__global__ __launch_bounds__(256) void k_h0_key(int64_t n, const double *h0, int32_t *key) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double h = h0[i];
  int k = H0_BINS - 1;
  if (h > 0.0 && h < INFINITY) {
    const double x = (log10(h) + 5.0) * (H0_BINS / (5.0 + log10(1200.0)));
    k = x < 0.0 ? 0 : (x >= (double)(H0_BINS - 1) ? H0_BINS - 1 : (int)x);
  }
  key[i] = k;
}
__global__ __launch_bounds__(256) void k_key_hist(int64_t n, const int32_t *key, int32_t *hist) {
  __shared__ int32_t cnt[H0_BINS];
  if (threadIdx.x < H0_BINS) cnt[threadIdx.x] = 0;
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) atomicAdd(&cnt[key[i]], 1);
  __syncthreads();
  if (threadIdx.x < H0_BINS) hist[(int64_t)blockIdx.x * H0_BINS + threadIdx.x] = cnt[threadIdx.x];
}
// In place: hist[b*K+k] <- start of block b's key-k cells (as k_class_scan).
__global__ __launch_bounds__(256) void k_key_scan(int64_t nb, int32_t *hist) {
  __shared__ int32_t tot[H0_BINS][256];
  __shared__ int32_t base[H0_BINS];
  const int64_t per = (nb + 255) / 256;
  const int64_t b0 = (int64_t)threadIdx.x * per;
  const int64_t b1 = b0 + per < nb ? b0 + per : nb;
  for (int k = 0; k < H0_BINS; ++k) {
    int32_t s = 0;
    for (int64_t b = b0; b < b1; ++b) s += hist[b * H0_BINS + k];
    tot[k][threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t run = 0;
    for (int k = 0; k < H0_BINS; ++k) {
      base[k] = run;
      for (int t = 0; t < 256; ++t) run += tot[k][t];
    }
  }
  __syncthreads();
  for (int k = 0; k < H0_BINS; ++k) {
    int32_t run = base[k];
    for (int t = 0; t < (int)threadIdx.x; ++t) run += tot[k][t];
    for (int64_t b = b0; b < b1; ++b) {
      int32_t c = hist[b * H0_BINS + k];
      hist[b * H0_BINS + k] = run;
      run += c;
    }
  }
}
__global__ __launch_bounds__(256) void k_key_scatter(int64_t n, const int32_t *key,
                                                     const int32_t *start, int32_t *order) {
  __shared__ int32_t ks[256];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int k = i < n ? key[i] : H0_BINS;
  ks[threadIdx.x] = k;
  __syncthreads();
  if (i < n) {
    int32_t r = 0;  // rank among the earlier cells of this block with the same key
    for (int t = 0; t < (int)threadIdx.x; ++t) r += ks[t] == k;
    order[start[(int64_t)blockIdx.x * H0_BINS + k] + r] = (int32_t)i;
  }
}

int64_t order_by_h0_scratch(int64_t n) { return n + (int64_t)H0_BINS * ((n + 255) / 256); }

void order_by_h0(int64_t n, const double *h0, int32_t *order, int32_t *scratch, hipStream_t s) {
  int64_t nb = (n + 255) / 256;
  int32_t *key = scratch, *hist = scratch + n;
  hipLaunchKernelGGL(k_h0_key, dim3((unsigned)nb), dim3(256), 0, s, n, h0, key);
  hipLaunchKernelGGL(k_key_hist, dim3((unsigned)nb), dim3(256), 0, s, n, key, hist);
  hipLaunchKernelGGL(k_key_scan, dim3(1), dim3(256), 0, s, nb, hist);
  hipLaunchKernelGGL(k_key_scatter, dim3((unsigned)nb), dim3(256), 0, s, n, key, hist, order);
}

}  // namespace gckpp
