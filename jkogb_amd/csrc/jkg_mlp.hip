// jkg_mlp.hip -- the learned step controller of the mixed solver: a 16-64-64-2
// MLP with BF16 weights and FP32 accumulation, its in-graph inference kernel
// and its training kernels (forward, backward, Adam).  No counterpart in the
// reference; the plain full-sort ordering is the controller's fallback.
//
// Parameter blob (FP32 master copy, 5378 floats, this order):
//   W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]
// The forward pass reads the weights rounded to BF16 (round to nearest even)
// and the biases in FP32; every product is accumulated in FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "jkg_mlp.hpp"
#include "jkg_step.hpp"

namespace jkg {

__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
  const uint32_t r = 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)((u + r) >> 16);
}
__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// ------------------------------------------------------------------ quantize
// q[i] = BF16(master[i]) for the weight matrices; biases are kept in FP32.
__global__ void k_mlp_quantize(const float *__restrict__ master, uint16_t *__restrict__ q) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MLP_NPARAM; i += gridDim.x * blockDim.x) {
    const bool bias = (i >= MLP_B1 && i < MLP_W2) || (i >= MLP_B2 && i < MLP_W3) || i >= MLP_B3;
    q[i] = bias ? 0 : f32_to_bf16(master[i]);
  }
}

// Weight (BF16-rounded) or bias (FP32) of the blob.
__device__ __forceinline__ float mlp_w(const uint16_t *q, int i) { return bf16_to_f32(q[i]); }

// ------------------------------------------------------------------ batched forward + backward
// One workgroup of 256 lanes per block of MLP_BS samples.  Activations live in
// LDS; every sum runs in a fixed order, so the partial gradients of a block
// (written to part[block][MLP_NPARAM]) are deterministic.  Loss: the weighted
// mean of 0.5 (y - t)^2 over the batch (wsum = sum of sample weights).
constexpr int MLP_BS = 64;

__global__ void __launch_bounds__(256) k_mlp_fwd_bwd(const uint16_t *__restrict__ q, const float *__restrict__ master,
                                                     const float *__restrict__ X, const float *__restrict__ T,
                                                     const float *__restrict__ SW, int ns, float inv_wsum,
                                                     float *__restrict__ part, float *__restrict__ Yout,
                                                     float *__restrict__ lpart) {
  __shared__ float xs[MLP_BS][MLP_IN];
  __shared__ float h1[MLP_BS][MLP_H + 1], h2[MLP_BS][MLP_H + 1];
  __shared__ float d1[MLP_BS][MLP_H + 1], d2[MLP_BS][MLP_H + 1];
  __shared__ float dy[MLP_BS][MLP_OUT];
  __shared__ float lsum[256];
  const int tid = threadIdx.x, s0 = blockIdx.x * MLP_BS;
  for (int e = tid; e < MLP_BS * MLP_IN; e += 256) {
    const int s = e / MLP_IN, i = e % MLP_IN;
    xs[s][i] = (s0 + s < ns) ? X[(size_t)(s0 + s) * MLP_IN + i] : 0.f;
  }
  __syncthreads();
  // layer 1 (pre-activations kept in h1, masks taken later from their sign)
  for (int e = tid; e < MLP_BS * MLP_H; e += 256) {
    const int s = e / MLP_H, j = e % MLP_H;
    float z = master[MLP_B1 + j];
    for (int i = 0; i < MLP_IN; ++i) z = __builtin_fmaf(mlp_w(q, MLP_W1 + j * MLP_IN + i), xs[s][i], z);
    h1[s][j] = z;
  }
  __syncthreads();
  for (int e = tid; e < MLP_BS * MLP_H; e += 256) {
    const int s = e / MLP_H, k = e % MLP_H;
    float z = master[MLP_B2 + k];
    for (int j = 0; j < MLP_H; ++j) {
      const float a = h1[s][j];
      z = __builtin_fmaf(mlp_w(q, MLP_W2 + k * MLP_H + j), a > 0.f ? a : 0.f, z);
    }
    h2[s][k] = z;
  }
  __syncthreads();
  float lacc = 0.f;
  if (tid < MLP_BS * MLP_OUT) {
    const int s = tid / MLP_OUT, o = tid % MLP_OUT;
    float z = master[MLP_B3 + o];
    for (int k = 0; k < MLP_H; ++k) {
      const float a = h2[s][k];
      z = __builtin_fmaf(mlp_w(q, MLP_W3 + o * MLP_H + k), a > 0.f ? a : 0.f, z);
    }
    float g = 0.f;
    if (s0 + s < ns) {
      const float w = SW ? SW[s0 + s] : 1.f;
      const float r = z - T[(size_t)(s0 + s) * MLP_OUT + o];
      g = w * r * inv_wsum;
      lacc = 0.5f * w * r * r * inv_wsum;
      if (Yout) Yout[(size_t)(s0 + s) * MLP_OUT + o] = z;
    }
    dy[s][o] = g;
  }
  lsum[tid] = lacc;
  __syncthreads();
  // dh2 -> d2 = dh2 * relu'(z2)
  for (int e = tid; e < MLP_BS * MLP_H; e += 256) {
    const int s = e / MLP_H, k = e % MLP_H;
    float g = 0.f;
    for (int o = 0; o < MLP_OUT; ++o) g = __builtin_fmaf(mlp_w(q, MLP_W3 + o * MLP_H + k), dy[s][o], g);
    d2[s][k] = h2[s][k] > 0.f ? g : 0.f;
  }
  __syncthreads();
  // dh1 -> d1 = dh1 * relu'(z1)
  for (int e = tid; e < MLP_BS * MLP_H; e += 256) {
    const int s = e / MLP_H, j = e % MLP_H;
    float g = 0.f;
    for (int k = 0; k < MLP_H; ++k) g = __builtin_fmaf(mlp_w(q, MLP_W2 + k * MLP_H + j), d2[s][k], g);
    d1[s][j] = h1[s][j] > 0.f ? g : 0.f;
  }
  __syncthreads();
  // parameter gradients of this block: sums over its samples in order
  float *pg = part + (size_t)blockIdx.x * MLP_NPARAM;
  for (int e = tid; e < MLP_NPARAM; e += 256) {
    float g = 0.f;
    if (e < MLP_B1) {  // W1[j][i]
      const int j = e / MLP_IN, i = e % MLP_IN;
      for (int s = 0; s < MLP_BS; ++s) g = __builtin_fmaf(d1[s][j], xs[s][i], g);
    } else if (e < MLP_W2) {  // b1
      const int j = e - MLP_B1;
      for (int s = 0; s < MLP_BS; ++s) g += d1[s][j];
    } else if (e < MLP_B2) {  // W2[k][j]
      const int k = (e - MLP_W2) / MLP_H, j = (e - MLP_W2) % MLP_H;
      for (int s = 0; s < MLP_BS; ++s) {
        const float a = h1[s][j];
        g = __builtin_fmaf(d2[s][k], a > 0.f ? a : 0.f, g);
      }
    } else if (e < MLP_W3) {  // b2
      const int k = e - MLP_B2;
      for (int s = 0; s < MLP_BS; ++s) g += d2[s][k];
    } else if (e < MLP_B3) {  // W3[o][k]
      const int o = (e - MLP_W3) / MLP_H, k = (e - MLP_W3) % MLP_H;
      for (int s = 0; s < MLP_BS; ++s) {
        const float a = h2[s][k];
        g = __builtin_fmaf(dy[s][o], a > 0.f ? a : 0.f, g);
      }
    } else {  // b3
      const int o = e - MLP_B3;
      for (int s = 0; s < MLP_BS; ++s) g += dy[s][o];
    }
    pg[e] = g;
  }
  // loss of this block (fixed-order tree)
  for (int o = 128; o > 0; o >>= 1) {
    __syncthreads();
    if (tid < o) lsum[tid] += lsum[tid + o];
  }
  if (tid == 0) lpart[blockIdx.x] = lsum[0];
}

// grad[e] = sum over blocks in order; loss = sum of the block losses.
__global__ void k_mlp_reduce(const float *__restrict__ part, int nblk, float *__restrict__ grad,
                             const float *__restrict__ lpart, float *__restrict__ loss) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < MLP_NPARAM) {
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += part[(size_t)b * MLP_NPARAM + e];
    grad[e] = g;
  }
  if (e == 0 && loss) {
    float l = 0.f;
    for (int b = 0; b < nblk; ++b) l += lpart[b];
    *loss = l;
  }
}

// Adam on the FP32 master copy (gradients of the BF16-rounded weights are
// applied to the master weights), then the BF16 copy is refreshed.
__global__ void k_mlp_adam(float *__restrict__ master, uint16_t *__restrict__ q, const float *__restrict__ grad,
                           float *__restrict__ m1, float *__restrict__ m2, float lr, float b1, float b2, float eps,
                           float c1, float c2, float wd) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_NPARAM) return;
  const bool bias = (e >= MLP_B1 && e < MLP_W2) || (e >= MLP_B2 && e < MLP_W3) || e >= MLP_B3;
  float g = grad[e];
  if (!bias) g = __builtin_fmaf(wd, master[e], g);
  const float a = __builtin_fmaf(b1, m1[e], (1.f - b1) * g);
  const float v = __builtin_fmaf(b2, m2[e], (1.f - b2) * g * g);
  m1[e] = a;
  m2[e] = v;
  const float w = master[e] - lr * (a / c1) / (__builtin_sqrtf(v / c2) + eps);
  master[e] = w;
  q[e] = bias ? 0 : f32_to_bf16(w);
}

// ------------------------------------------------------------------ the controller
// One wave, at the start of each step of the mixed solver, sets this step's
// N_2 (st->n2, a fraction of the problem's N_2) and threshold (st->tau) from
// the features the previous step left (st->feat):
//   mode 1: the MLP; outputs y clamped to [0, 1] map to
//           N_2 fraction = 0.75 + 0.25 y0 and log10 tau = -16 + 10 y1;
//   mode 2: a phase table t[8]: phases by log10(TW / TW_0) against t[0] > t[1],
//           actions (N_2 fraction, log10 tau) in t[2..7].
// The plain full ordering (all of N_2, no threshold) is the fallback: on the
// first step, once TW has fallen below 10^lhandoff TW_0 (st->lhandoff, 1e-10 by
// default), and for outputs that are not finite.  log10 tau <= -15.5 means no threshold.
__global__ void __launch_bounds__(64) k_controller(StepState *st, int mode, const float *__restrict__ master,
                                                   const uint16_t *__restrict__ q, const float *__restrict__ table) {
  __shared__ float h1[MLP_H], h2[MLP_H];
  if (!st->active) return;
  const int lane = threadIdx.x;
  const float *f = st->feat;
  const bool first = (st->step == 0);
  const float lrel = f[1] * 20.f;  // log10(TW / TW_0) after the last step
  float n2f = 1.f, ltau = -16.f;
  if (mode == 1 && !first) {
    float z = master[MLP_B1 + lane];
    for (int i = 0; i < MLP_IN; ++i) z = __builtin_fmaf(mlp_w(q, MLP_W1 + lane * MLP_IN + i), f[i], z);
    h1[lane] = z > 0.f ? z : 0.f;
    __syncthreads();
    z = master[MLP_B2 + lane];
    for (int j = 0; j < MLP_H; ++j) z = __builtin_fmaf(mlp_w(q, MLP_W2 + lane * MLP_H + j), h1[j], z);
    h2[lane] = z > 0.f ? z : 0.f;
    __syncthreads();
    float y[MLP_OUT];
    for (int o = 0; o < MLP_OUT; ++o) {
      float a = mlp_w(q, MLP_W3 + o * MLP_H + lane) * h2[lane];
      for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);  // same tree in every lane
      y[o] = a + master[MLP_B3 + o];
    }
    if (y[0] == y[0] && y[1] == y[1] && lrel > st->lhandoff) {
      const float y0 = y[0] < 0.f ? 0.f : (y[0] > 1.f ? 1.f : y[0]);
      const float y1 = y[1] < 0.f ? 0.f : (y[1] > 1.f ? 1.f : y[1]);
      n2f = 0.75f + 0.25f * y0;
      ltau = -16.f + 10.f * y1;
    }
  } else if (mode == 2) {
    const int ph = first ? 0 : (lrel > table[0] ? 0 : (lrel > table[1] ? 1 : 2));
    n2f = table[2 + 2 * ph];
    ltau = table[3 + 2 * ph];
    if (!first && lrel <= st->lhandoff) {
      n2f = 1.f;
      ltau = -16.f;
    }
    if (lane == 0) st->ctl_phase = ph;
  }
  if (lane == 0) {
    const int m = st->n2max;
    int n2 = (int)(n2f * (float)m + 0.5f);
    st->n2 = n2 < 1 ? 1 : (n2 > m ? m : n2);
    if (ltau > st->ltau_cap) ltau = st->ltau_cap;  // the caller's bound on the threshold
    st->tau = (ltau <= -15.5f) ? 0.0 : exp10((double)ltau);
    st->n2f = n2f;
    st->ltau = ltau;
  }
}

void controller_launch(DeviceStep &d, hipStream_t s) {
  hipLaunchKernelGGL(k_controller, dim3(1), dim3(64), 0, s, d.st, d.ctl_mode, d.ctl_master, d.ctl_q, d.ctl_table);
}

// ------------------------------------------------------------------ host launchers
void mlp_quantize_launch(const float *master, uint16_t *q, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_quantize, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, master, q);
}

int mlp_blocks(int ns) { return (ns + MLP_BS - 1) / MLP_BS; }

void mlp_grad_launch(const uint16_t *q, const float *master, const float *X, const float *T, const float *SW,
                     int ns, float inv_wsum, float *part, float *lpart, float *grad, float *loss, float *Yout,
                     hipStream_t s) {
  const int nb = mlp_blocks(ns);
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(256), 0, s, q, master, X, T, SW, ns, inv_wsum, part, Yout,
                     lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, part, nb, grad, lpart, loss);
}

void mlp_adam_launch(float *master, uint16_t *q, const float *grad, float *m1, float *m2, float lr, int t,
                     float wd, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = 1.f, c2 = 1.f;
  for (int i = 0; i < t; ++i) {
    c1 *= b1;
    c2 *= b2;
  }
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, master, q, grad, m1, m2, lr, b1,
                     b2, 1e-8f, 1.f - c1, 1.f - c2, wd);
}

}  // namespace jkg
