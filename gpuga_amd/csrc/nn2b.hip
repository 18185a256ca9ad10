// NN2B population path.
//
// U(r) of NN2B depends on r only (src/nn2b.cu:49-78), and every pair of the
// training set is evaluated by every individual.  Instead of running the
// 1 -> n -> n -> 1 network once per (individual, pair), each generation first
// tabulates, for every individual, the network output p(r) and its first two
// derivatives on a uniform grid covering the training set's pair distances
// (this is where the n x n hidden layer runs, on MFMA or FP32 VALU), then the
// pair kernel reconstructs p and p' at each pair distance with cubic Hermite
// interpolation (p from p, p'; p' from p', p'') and applies the reference's
// cutoff taper and force quirk f2 = (p' fc + (p fc) fc') / r
// (src/nn2b.cu:143-148).
#include "kernels.h"
#include "population.h"

namespace gpuga {

namespace {

// table layout: [K][G][3][64] (p, p', p'' for the 64 individuals of block k)
__device__ inline size_t tab_index(int k, int G, int g, int q, int lane)
{
  return (((size_t)k * G + g) * 3 + q) * WAVE + lane;
}

// FP32 tabulation: block per individual (weights staged in LDS), thread per
// grid point.  NMAX bounds the neuron count for register-resident layers.
template <int NMAX>
__global__ void __launch_bounds__(256) nn2b_tabulate_f32(PopCtl ctl, NN2BGrid grid,
                                                         float* __restrict__ table)
{
  extern __shared__ float w[];
  const int p = blockIdx.y;
  const int n = grid.n, V = ctl.V;
  for (int v = threadIdx.x; v < V; v += blockDim.x)
    w[v] = denormalise(ctl.pmin[v], ctl.pmax[v], ctl.genes[(size_t)p * V + v]);
  __syncthreads();
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= grid.G) return;
  const float* w0 = w;
  const float* b0 = w + n;
  const float* w1 = w + 2 * n;
  const float* b1 = w + n * (n + 2);
  const float* w2 = w + n * (n + 3);
  const float b2 = w[n * (n + 4)];
  const float d = grid.d0 + grid.h * g;
  float x1[NMAX], y1[NMAX], y2[NMAX];
#pragma unroll
  for (int m = 0; m < NMAX; ++m) {
    if (m < n) {
      const float t = tanhf(w0[m] * d - b0[m]);
      const float s = 1.0f - t * t;
      x1[m] = t;
      y1[m] = s * w0[m];                      // d x1 / dr
      y2[m] = -2.0f * t * s * w0[m] * w0[m];  // d2 x1 / dr2
    }
  }
  float pv = 0.0f, pd = 0.0f, pdd = 0.0f;
  for (int a = 0; a < n; ++a) {
    float z = 0.0f, zd = 0.0f, zdd = 0.0f;
#pragma unroll
    for (int m = 0; m < NMAX; ++m) {
      if (m < n) {
        const float wm = w1[a * n + m];
        z += wm * x1[m];
        zd += wm * y1[m];
        zdd += wm * y2[m];
      }
    }
    const float x2 = tanhf(z - b1[a]);
    const float s = 1.0f - x2 * x2;
    pv += w2[a] * x2;
    pd += w2[a] * s * zd;
    pdd += w2[a] * (s * zdd - 2.0f * x2 * s * zd * zd);
  }
  const int k = p / WAVE, lane = p % WAVE;
  table[tab_index(k, grid.G, g, 0, lane)] = pv - b2;
  table[tab_index(k, grid.G, g, 1, lane)] = pd;
  table[tab_index(k, grid.G, g, 2, lane)] = pdd;
}

} // namespace

struct NN2BPop {
  static constexpr bool kForcePass = false;
  struct Lane {
    int dummy;
  };
  struct Extra {
    NN2BGrid g;
    const float* table;
  };

  __device__ static Lane make(const PopCtl&, int) { return Lane{0}; }

  // p(r) and p'(r) of this lane's individual by cubic Hermite interpolation
  __device__ static void eval(const Extra& ex, const float* __restrict__ tk, float d, float& p,
                              float& pd)
  {
    float u = (d - ex.g.d0) / ex.g.h;
    int g = (int)u;
    g = max(0, min(g, ex.g.G - 2));
    const float t = u - (float)g;
    const float t2 = t * t, t3 = t2 * t;
    const float h00 = 2.0f * t3 - 3.0f * t2 + 1.0f, h10 = t3 - 2.0f * t2 + t;
    const float h01 = -2.0f * t3 + 3.0f * t2, h11 = t3 - t2;
    const float* r0 = tk + (size_t)g * 3 * WAVE;
    const float* r1 = r0 + 3 * WAVE;
    const float p0 = r0[0], d0 = r0[WAVE], s0 = r0[2 * WAVE];
    const float p1 = r1[0], d1 = r1[WAVE], s1 = r1[2 * WAVE];
    const float h = ex.g.h;
    p = h00 * p0 + h10 * h * d0 + h01 * p1 + h11 * h * d1;
    pd = h00 * d0 + h10 * h * s0 + h01 * d1 + h11 * h * s1;
  }

  __device__ static void taper(float d, float& fc, float& fcp)
  {
    if (d < NN2B_R1) {
      fc = 1.0f;
      fcp = 0.0f;
    } else if (d < NN2B_R2) {
      float sn, cs;
      sincosf(NN2B_PI_FACTOR * (d - NN2B_R1), &sn, &cs);
      fc = 0.5f * cs + 0.5f;
      fcp = -sn * NN2B_PI_FACTOR * 0.5f;
    } else {
      fc = 0.0f;
      fcp = 0.0f;
    }
  }

  __device__ static void chunk(const DevSet& s, const Lane&, const WorkItem& it, const PopCtl&,
                               const Extra& ex, Acc& acc)
  {
    const int lane = __lane_id();
    const float* tk = ex.table + (size_t)it.k * ex.g.G * 3 * WAVE + lane;
    const size_t N = s.N;
    if (it.c < s.Nc_force) {
      for (int i = it.a0; i < it.a1; ++i) {
        const int nn = s.nn[i], b0 = s.bond0[i];
        float fx = 0, fy = 0, fz = 0;
        for (int u = 0; u < nn; ++u) {
          const float4 r = s.bond[b0 + u];
          float p, pd, fc, fcp;
          eval(ex, tk, r.w, p, pd);
          taper(r.w, fc, fcp);
          const float p2 = p * fc;
          const float f2 = (pd * fc + p2 * fcp) / r.w;
          fx += r.x * f2;
          fy += r.y * f2;
          fz += r.z * f2;
        }
        const float dx = fx - s.force_ref[i], dy = fy - s.force_ref[i + N],
                    dz = fz - s.force_ref[i + 2 * N];
        acc.v[7] += (double)dx * dx + (double)dy * dy + (double)dz * dz;
      }
    } else {
      for (int i = it.a0; i < it.a1; ++i) {
        const int nn = s.nn[i], b0 = s.bond0[i];
        float e = 0, wxx = 0, wyy = 0, wzz = 0, wxy = 0, wyz = 0, wzx = 0;
        for (int u = 0; u < nn; ++u) {
          const int b = b0 + u;
          if (s.nl[b] < i) continue; // each unordered pair once, full weight
          const float4 r = s.bond[b];
          float p, pd, fc, fcp;
          eval(ex, tk, r.w, p, pd);
          taper(r.w, fc, fcp);
          const float p2 = p * fc;
          const float f2 = (pd * fc + p2 * fcp) / r.w;
          const float rx = r.x * f2, ry = r.y * f2, rz = r.z * f2;
          wxx -= r.x * rx;
          wyy -= r.y * ry;
          wzz -= r.z * rz;
          wxy -= r.x * ry;
          wyz -= r.y * rz;
          wzx -= r.z * rx;
          e += p2;
        }
        acc.v[0] += e;
        acc.v[1] += wxx;
        acc.v[2] += wyy;
        acc.v[3] += wzz;
        acc.v[4] += wxy;
        acc.v[5] += wyz;
        acc.v[6] += wzx;
      }
    }
  }

  __device__ static double force_pass(const DevSet&, const PopCtl&, int, int, int, int)
  {
    return 0.0;
  }
};

void launch_nn2b_tabulate_f32(const PopCtl& ctl, const NN2BGrid& g, float* table, hipStream_t st)
{
  const dim3 block(256), grid((g.G + 255) / 256, ctl.P);
  const size_t lds = sizeof(float) * ctl.V;
  if (g.n <= 8)
    nn2b_tabulate_f32<8><<<grid, block, lds, st>>>(ctl, g, table);
  else if (g.n <= 16)
    nn2b_tabulate_f32<16><<<grid, block, lds, st>>>(ctl, g, table);
  else if (g.n <= 32)
    nn2b_tabulate_f32<32><<<grid, block, lds, st>>>(ctl, g, table);
  else if (g.n <= 64)
    nn2b_tabulate_f32<64><<<grid, block, lds, st>>>(ctl, g, table);
  else
    fail("NN2B supports at most 64 neurons");
  GPUGA_HIP(hipGetLastError());
}

void launch_nn2b_tabulate_bf16(const PopCtl& ctl, const NN2BGrid& g, float* table, hipStream_t st);

void launch_nn2b_tabulate(const PopCtl& ctl, const NN2BGrid& g, float* table, hipStream_t st)
{
  if (g.bf16)
    launch_nn2b_tabulate_bf16(ctl, g, table, st);
  else
    launch_nn2b_tabulate_f32(ctl, g, table, st);
}

void launch_nn2b_population(const PopCtl& ctl, const DevSet& s, const NN2BGrid& g,
                            const float* table, int grid, int block, hipStream_t st)
{
  NN2BPop::Extra ex{g, table};
  population_kernel<NN2BPop><<<grid, block, 0, st>>>(ctl, s, ex);
  GPUGA_HIP(hipGetLastError());
}

int nn2b_population_occupancy(int block)
{
  int n = 0;
  GPUGA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, population_kernel<NN2BPop>, block, 0));
  return n;
}

} // namespace gpuga
