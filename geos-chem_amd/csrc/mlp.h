// Sizes, parameter layout and launchers of the step-size MLP (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gckpp {

constexpr int MLP_NJ = 12;                  // photolysis channels used
constexpr int MLP_NS = 24;                  // species used
constexpr int MLP_NIN = 4 + MLP_NJ + MLP_NS + 2;  // 42 inputs
constexpr int MLP_F = 48;                   // inputs padded to the MFMA K step
constexpr int MLP_H = 128;                  // hidden units per layer (3 layers)
constexpr int MLP_O = 16;                   // outputs padded (log h0 + 4 logits used)
// fp32 parameters, W[in][out] row-major then bias, layer after layer
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_F * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_H;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_H;
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_H * MLP_O;
constexpr int MLP_NPARAM = MLP_OFF_B4 + MLP_O;
// bf16 packed weights (MFMA A-operand tiles), in units of 2 bytes
constexpr int MLP_NPACK = (MLP_F * MLP_H + 2 * MLP_H * MLP_H + MLP_H * MLP_O);

// Inputs: feature k = (x_k - mean[k]) * rstd[k] of
//   0 TEMP, 1 log10 PRESS, 2 log10 NUMDEN, 3 SUNCOS,
//   4..15 log10(PHOTOL(jslot)+1e-20), 16..39 log10(C(spc)+1),
//   40 ln(hprev) and 41 = 1 when the cell has a previous step size hprev > 0
//   (KPPHvalue), else 0 and 0; 42..47 zero.
// Output 0 is (ln h0 - out_mu) / out_sigma.
struct MlpDesc {
  int32_t jslot[MLP_NJ];
  int32_t spc[MLP_NS];
  float mean[MLP_F], rstd[MLP_F];
  float out_mu, out_sigma;
};

void mlp_pack(const float *params, unsigned short *pk, hipStream_t s);
void mlp_features(const MlpDesc &d, int64_t n, const double *met, const double *photol,
                  const double *conc, const double *hprev, float *X, hipStream_t s);
void mlp_predict(const MlpDesc &d, int64_t n, const double *met, const double *photol,
                 const double *conc, const double *hprev, const unsigned short *pk,
                 const float *params, double safety, double hmin, double hmax, double *h0,
                 int32_t *cls, hipStream_t s);
void mlp_forward_f32(const float *P, int64_t nb, const float *X, float *act, hipStream_t s);
void mlp_grad_f32(const float *P, int64_t nb, const float *X, const float *yh, const int32_t *ycls,
                  float wh, float wc, float *act, float *dO, float *dA, float *G, float *lossp,
                  hipStream_t s);
// order[0..n) = cells sorted by class, class 3 first, stable; scratch: 4*ceil(n/256) ints
void order_by_class(int64_t n, const int32_t *cls, int32_t *order, int32_t *scratch, hipStream_t s);
// order[0..n) = cells sorted by predicted h0, smallest first (32 bins of
// log10 h0), stable; scratch: order_by_h0_scratch(n) ints
void order_by_h0(int64_t n, const double *h0, int32_t *order, int32_t *scratch, hipStream_t s);
int64_t order_by_h0_scratch(int64_t n);
void mlp_sum(const float *x, int64_t n, float *out, hipStream_t s);  // out[0] = sum x[0..n)
void mlp_adam(float *P, const float *G, float *m, float *v, float lr, int step, hipStream_t s);

}  // namespace gckpp
