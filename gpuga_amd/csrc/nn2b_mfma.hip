// NN2B tabulation with the n x n hidden layer on BF16 MFMA (FP32 accumulate).
//
// One wavefront per (individual, group of 32-point grid tiles).  For a tile,
// the first hidden layer and its r-derivatives (x1, dx1/dr, d2x1/dr2;
// src/nn2b.cu:52-56 and their derivatives) form three 32 x 32 A operands
// (grid point x input neuron); the layer-2 weights W1 (src/nn2b.cu:62-66) are
// the B operand; v_mfma_f32_32x32x8_bf16 produces z, dz/dr, d2z/dr2 for the
// 32 x 32 (grid point x output neuron) tile.  The epilogue applies tanh and
// the chain rule per output neuron and folds the output layer
// (src/nn2b.cu:71-77) through LDS in a fixed order.  n <= 32 (zero padded).
#include "kernels.h"

namespace gpuga {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bfloat16, round to nearest even
__device__ inline short to_bf16(float f)
{
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

constexpr int TILE = 32;

__global__ void __launch_bounds__(64) nn2b_tabulate_bf16(PopCtl ctl, NN2BGrid grid, int tiles,
                                                         float* __restrict__ table)
{
  __shared__ float red[3][TILE][TILE + 1];
  const int p = blockIdx.y;
  const int l = threadIdx.x, col = l & 31, kg = l >> 5;
  const int n = grid.n, V = ctl.V;
  const float* gp = ctl.genes + (size_t)p * V;
  auto W = [&](int v) { return denormalise(ctl.pmin[v], ctl.pmax[v], gp[v]); };

  // A-operand neurons of this lane: m = 8 s + 4 kg + j
  float w0m[16], b0m[16];
  bf16x4 B[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = 8 * s + 4 * kg + j;
      const bool ok = m < n;
      w0m[4 * s + j] = ok ? W(m) : 0.0f;
      b0m[4 * s + j] = ok ? W(n + m) : 0.0f;
      B[s][j] = to_bf16(ok && col < n ? W(2 * n + col * n + m) : 0.0f);
    }
  }
  const float b1c = col < n ? W(n * (n + 2) + col) : 0.0f;
  const float w2c = col < n ? W(n * (n + 3) + col) : 0.0f;
  const float b2 = W(n * (n + 4));
  const int k = p / WAVE, lp = p % WAVE;

  for (int tt = 0; tt < tiles; ++tt) {
    const int g0 = (blockIdx.x * tiles + tt) * TILE;
    if (g0 >= grid.G) break;
    const float d = grid.d0 + grid.h * (float)(g0 + col);
    bf16x4 AX[4], AY[4], AZ[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float w = w0m[4 * s + j];
        const float t = tanhf(w * d - b0m[4 * s + j]);
        const float sg = 1.0f - t * t;
        AX[s][j] = to_bf16(t);
        AY[s][j] = to_bf16(sg * w);
        AZ[s][j] = to_bf16(-2.0f * t * sg * w * w);
      }
    }
    f32x16 Z = {0}, Zd = {0}, Zdd = {0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      Z = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(AX[s], B[s], Z, 0, 0, 0);
      Zd = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(AY[s], B[s], Zd, 0, 0, 0);
      Zdd = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(AZ[s], B[s], Zdd, 0, 0, 0);
    }
    // this lane holds output neuron `col` for grid rows 8 (r/4) + 4 kg + r%4
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r >> 2) + 4 * kg + (r & 3);
      const float x2 = tanhf(Z[r] - b1c);
      const float s = 1.0f - x2 * x2;
      red[0][row][col] = w2c * x2;
      red[1][row][col] = w2c * s * Zd[r];
      red[2][row][col] = w2c * (s * Zdd[r] - 2.0f * x2 * s * Zd[r] * Zd[r]);
    }
    __syncthreads();
    for (int idx = l; idx < 3 * TILE; idx += WAVE) {
      const int q = idx / TILE, row = idx % TILE;
      float sum = 0.0f;
      for (int c = 0; c < TILE; ++c) sum += red[q][row][c];
      if (q == 0) sum -= b2;
      const int g = g0 + row;
      if (g < grid.G) table[(((size_t)k * grid.G + g) * 3 + q) * WAVE + lp] = sum;
    }
    __syncthreads();
  }
}

} // namespace

void launch_nn2b_tabulate_f32(const PopCtl& ctl, const NN2BGrid& g, float* table, hipStream_t st);

void launch_nn2b_tabulate_bf16(const PopCtl& ctl, const NN2BGrid& g, float* table, hipStream_t st)
{
  if (g.n > 32) {
    launch_nn2b_tabulate_f32(ctl, g, table, st);
    return;
  }
  const int ntiles = (g.G + TILE - 1) / TILE;
  const int tiles = 4;
  const dim3 grid((ntiles + tiles - 1) / tiles, ctl.P), block(64);
  nn2b_tabulate_bf16<<<grid, block, 0, st>>>(ctl, g, tiles, table);
  GPUGA_HIP(hipGetLastError());
}

} // namespace gpuga
