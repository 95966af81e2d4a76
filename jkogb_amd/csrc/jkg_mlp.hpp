// jkg_mlp.hpp -- layout of the step controller's MLP (16-64-64-2).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jkg {

constexpr int MLP_IN = 16, MLP_H = 64, MLP_OUT = 2;
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_H * MLP_IN;
constexpr int MLP_W2 = MLP_B1 + MLP_H;
constexpr int MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H;
constexpr int MLP_B3 = MLP_W3 + MLP_OUT * MLP_H;
constexpr int MLP_NPARAM = MLP_B3 + MLP_OUT;  // 5378

void mlp_quantize_launch(const float *master, uint16_t *q, hipStream_t s);
int mlp_blocks(int ns);
void mlp_grad_launch(const uint16_t *q, const float *master, const float *X, const float *T, const float *SW,
                     int ns, float inv_wsum, float *part, float *lpart, float *grad, float *loss, float *Yout,
                     hipStream_t s);
void mlp_adam_launch(float *master, uint16_t *q, const float *grad, float *m1, float *m2, float lr, int t, float wd,
                     hipStream_t s);

}  // namespace jkg
