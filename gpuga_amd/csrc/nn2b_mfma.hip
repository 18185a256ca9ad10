// NN2B tabulation with the n x n hidden layer on BF16 MFMA (FP32 accumulate).
//
// One wavefront per (individual, group of 32-point grid tiles).  For a tile,
// the first hidden layer and its r-derivatives (x1, dx1/dr, d2x1/dr2;
// src/nn2b.cu:52-56 and their derivatives) form three 32 x 32 A operands
// (grid point x input neuron); the layer-2 weights W1 (src/nn2b.cu:62-66) are
// the B operand; v_mfma_f32_32x32x16_bf16 produces z, dz/dr, d2z/dr2 for the
// 32 x 32 (grid point x output neuron) tile.  Every operand is split into two
// BF16 parts, x = hi + lo with hi = bf16(x) and lo = bf16(x - hi) (16
// significant bits together), and a product is accumulated as
// lo_A hi_B + hi_A lo_B + hi_A hi_B (the lo_A lo_B term, 2^-16 relative, is
// dropped), so the layer keeps nearly FP32 accuracy on BF16 inputs.  A lane
// holds the same input neurons in its A and B operands, so the products pair
// up whatever order the instruction gives the K index inside a lane.  The
// epilogue applies tanh and the chain rule per output neuron and folds the
// output layer (src/nn2b.cu:71-77) through LDS in a fixed order.  n <= 32
// (zero padded).
#include "kernels.h"

namespace gpuga {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// x = hi + lo, both round-to-nearest-even BF16, into element j of two operands
__device__ inline void split_bf16(float x, bf16x8& hi, bf16x8& lo, int j)
{
  const __bf16 h = (__bf16)x;
  hi[j] = h;
  lo[j] = (__bf16)(x - (float)h);
}

// C += A B with split operands (smallest terms first)
__device__ inline f32x16 mfma3(const bf16x8& ah, const bf16x8& al, const bf16x8& bh,
                               const bf16x8& bl, f32x16 c)
{
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, c, 0, 0, 0);
}

constexpr int TILE = 32;

__global__ void __launch_bounds__(64) nn2b_tabulate_bf16(PopCtl ctl, NN2BGrid grid, int tiles,
                                                         float* __restrict__ table)
{
  __shared__ float red[3][TILE][TILE + 1];
  const int p = blockIdx.y, code = blockIdx.z;
  const int l = threadIdx.x, col = l & 31, kg = l >> 5;
  const int n = grid.n, V = ctl.V;
  const float* gp = ctl.genes + (size_t)p * V;
  auto W = [&](int v) { return denormalise(ctl.pmin[v], ctl.pmax[v], gp[v]); };

  // input neurons of this lane, in its A and B operands alike:
  // m = 16 s + 8 kg + j (K step s, element j)
  float w0m[16], b0m[16];
  bf16x8 Bh[2], Bl[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int m = 16 * s + 8 * kg + j;
      const bool ok = m < n;
      w0m[8 * s + j] = ok ? W(m) : 0.0f;
      // the species-pair input (codes > 1) enters as a shift of the bias
      b0m[8 * s + j] =
        ok ? W(n + m) - (grid.codes > 1 ? W(n * (n + 4) + 1 + m) * (float)code : 0.0f) : 0.0f;
      split_bf16(ok && col < n ? W(2 * n + col * n + m) : 0.0f, Bh[s], Bl[s], j);
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
    f32x16 Z = {0}, Zd = {0}, Zdd = {0};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      bf16x8 AXh, AXl, AYh, AYl, AZh, AZl;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float w = w0m[8 * s + j];
        const float t = ftanh(w * d - b0m[8 * s + j]);
        const float sg = 1.0f - t * t;
        split_bf16(t, AXh, AXl, j);
        split_bf16(sg * w, AYh, AYl, j);
        split_bf16(-2.0f * t * sg * w * w, AZh, AZl, j);
      }
      Z = mfma3(AXh, AXl, Bh[s], Bl[s], Z);
      Zd = mfma3(AYh, AYl, Bh[s], Bl[s], Zd);
      Zdd = mfma3(AZh, AZl, Bh[s], Bl[s], Zdd);
    }
    // this lane holds output neuron `col` for grid rows 8 (r/4) + 4 kg + r%4
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r >> 2) + 4 * kg + (r & 3);
      const float x2 = ftanh(Z[r] - b1c);
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
      if (g < grid.G)
        table[((((size_t)code * ctl.K + k) * grid.G + g) * 3 + q) * WAVE + lp] = sum;
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
  const dim3 grid((ntiles + tiles - 1) / tiles, ctl.P, g.codes), block(64);
  nn2b_tabulate_bf16<<<grid, block, 0, st>>>(ctl, g, tiles, table);
  GPUGA_HIP(hipGetLastError());
}

} // namespace gpuga
