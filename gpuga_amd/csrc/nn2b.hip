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

// table layout: [code][K][G][3][64] (p, p', p'' for the 64 individuals of
// block k, per species code of the pair; one code without the species input)
__device__ inline size_t tab_index(int code, int K, int k, int G, int g, int q, int lane)
{
  return ((((size_t)code * K + k) * G + g) * 3 + q) * WAVE + lane;
}

// FP32 tabulation: block per individual (weights staged in LDS), thread per
// grid point.  NMAX bounds the neuron count for register-resident layers.
template <int NMAX>
__global__ void __launch_bounds__(256) nn2b_tabulate_f32(PopCtl ctl, NN2BGrid grid,
                                                         float* __restrict__ table)
{
  extern __shared__ float w[];
  const int p = blockIdx.y, code = blockIdx.z;
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
  const float* ws = w + n * (n + 4) + 1; // species-pair weights (codes > 1)
  const float d = grid.d0 + grid.h * g;
  float x1[NMAX], y1[NMAX], y2[NMAX];
#pragma unroll
  for (int m = 0; m < NMAX; ++m) {
    if (m < n) {
      // the species-pair input as a shift of the bias (none for one code)
      const float bm = grid.codes > 1 ? b0[m] - ws[m] * (float)code : b0[m];
      const float t = ftanh(w0[m] * d - bm);
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
    const float x2 = ftanh(z - b1[a]);
    const float s = 1.0f - x2 * x2;
    pv += w2[a] * x2;
    pd += w2[a] * s * zd;
    pdd += w2[a] * (s * zdd - 2.0f * x2 * s * zd * zd);
  }
  const int k = p / WAVE, lane = p % WAVE;
  table[tab_index(code, ctl.K, k, grid.G, g, 0, lane)] = pv - b2;
  table[tab_index(code, ctl.K, k, grid.G, g, 1, lane)] = pd;
  table[tab_index(code, ctl.K, k, grid.G, g, 2, lane)] = pdd;
}

} // namespace

struct NN2BPop {
  static constexpr bool kForcePass = false;
  static constexpr int kFoldValues = 12;
  static constexpr bool kPacked = false;
  struct Lane {
    int dummy;
  };
  struct Extra {
    NN2BGrid g;
    const float* table;
    long code_stride; // floats from one species code's tables to the next
  };

  __device__ static Lane make(const PopCtl&, int) { return Lane{0}; }

  // One pair from its table entry u = (unit vector, d), w = Hermite weights,
  // c = (fc, fc', grid cell): p = U fc and G = d f2 = U' fc + (U fc) fc'
  // (the reference's force expression, src/nn2b.cu:143-148), with U and U'
  // of this lane's individual interpolated from the rows (p, p', p'') of
  // the cell's two grid points.
  struct Rows {
    float p0, d0, s0, p1, d1, s1;
  };
  __device__ static __forceinline__ Rows rows(const float* __restrict__ tk, long code_stride,
                                              const float4& c)
  {
    const float* r0 = tk + (long)__float_as_int(c.w) * code_stride +
                      (size_t)__float_as_int(c.z) * 3 * WAVE;
    return Rows{r0[0], r0[WAVE], r0[2 * WAVE], r0[3 * WAVE], r0[4 * WAVE], r0[5 * WAVE]};
  }
  __device__ static __forceinline__ void pair(const Rows& r, const float4& w, const float4& c,
                                              float& p2, float& G)
  {
    const float p = fmaf(w.x, r.p0, fmaf(w.y, r.d0, fmaf(w.z, r.p1, w.w * r.d1)));
    const float pd = fmaf(w.x, r.d0, fmaf(w.y, r.s0, fmaf(w.z, r.d1, w.w * r.s1)));
    p2 = p * c.x;
    G = fmaf(pd, c.x, p2 * c.y);
  }
  // The pairs of an atom's table, four (then two, then one) at a time: the
  // pairs' table rows are all loaded before any is evaluated, so the
  // lookups (a gather that depends on the scalar load of the cell) are
  // waited for once per step rather than once per pair.  R float4 per pair,
  // the last two the Hermite weights and the cell; f(entry, rows, weights,
  // cell) reads the rest of the entry itself.
  template <int R, class F>
  __device__ static __forceinline__ void each_pair(const float* __restrict__ tk, long cs,
                                                   const float4* tab, int n, F&& f)
  {
    int k = 0;
    for (; k + 4 <= n; k += 4) {
      const float4 c0 = ldu(tab, R * k + R - 1), c1 = ldu(tab, R * (k + 1) + R - 1);
      const float4 c2 = ldu(tab, R * (k + 2) + R - 1), c3 = ldu(tab, R * (k + 3) + R - 1);
      const Rows r0 = rows(tk, cs, c0), r1 = rows(tk, cs, c1), r2 = rows(tk, cs, c2),
                 r3 = rows(tk, cs, c3);
      f(tab + R * k, r0, ldu(tab, R * k + R - 2), c0);
      f(tab + R * (k + 1), r1, ldu(tab, R * (k + 1) + R - 2), c1);
      f(tab + R * (k + 2), r2, ldu(tab, R * (k + 2) + R - 2), c2);
      f(tab + R * (k + 3), r3, ldu(tab, R * (k + 3) + R - 2), c3);
    }
    for (; k + 2 <= n; k += 2) {
      const float4 c0 = ldu(tab, R * k + R - 1), c1 = ldu(tab, R * (k + 1) + R - 1);
      const float4 w0 = ldu(tab, R * k + R - 2), w1 = ldu(tab, R * (k + 1) + R - 2);
      const Rows r0 = rows(tk, cs, c0), r1 = rows(tk, cs, c1);
      f(tab + R * k, r0, w0, c0);
      f(tab + R * (k + 1), r1, w1, c1);
    }
    if (k < n) {
      const float4 c = ldu(tab, R * k + R - 1), w = ldu(tab, R * k + R - 2);
      f(tab + R * k, rows(tk, cs, c), w, c);
    }
  }

  static constexpr bool kHeaderAhead = false;
  __device__ static void chunk(const DevSet& s, const Lane&, const WorkItem& it, const PopCtl&,
                               const Extra& ex, Acc& acc)
  {
    const int lane = __lane_id();
    const float* tk = ex.table + (size_t)it.k * ex.g.G * 3 * WAVE + lane;
    const long N = s.N;
    if (it.c < s.Nc_force) {
      for (int i = it.a0; i < it.a1; ++i) {
        const int4 h = ldu(s.pair_hdr, i);
        const float4* tab = s.pair_tab + (long)h.x;
        float fx = 0, fy = 0, fz = 0;
        // a force configuration's pair: (unit vector, d), weights, cell
        each_pair<3>(tk, ex.code_stride, tab, h.y,
                     [&](const float4* e, const Rows& r, const float4& w, const float4& c) {
          const float4 u = ldu(e, 0);
          float p2, G;
          pair(r, w, c, p2, G);
          fx = fmaf(u.x, G, fx);
          fy = fmaf(u.y, G, fy);
          fz = fmaf(u.z, G, fz);
        });
        const float dx = fx - ldu(s.force_ref, i), dy = fy - ldu(s.force_ref, i + N),
                    dz = fz - ldu(s.force_ref, i + 2 * N);
        acc.v[7] += (double)dx * dx + (double)dy * dy + (double)dz * dz;
      }
    } else {
      // each unordered pair once (the table holds j > i), full weight
      for (int i = it.a0; i < it.a1; ++i) {
        const int4 h = ldu(s.pair_hdr, i);
        const float4* tab = s.pair_tab + (long)h.x;
        float e = 0, wxx = 0, wyy = 0, wzz = 0, wxy = 0, wyz = 0, wzx = 0;
        // an energy configuration's pair: its virial coefficients u_a u_b d
        // (xx, yy, zz, xy; yz, zx), weights, cell: one FMA per component
        each_pair<4>(tk, ex.code_stride, tab, h.y,
                     [&](const float4* en, const Rows& r, const float4& w, const float4& c) {
          const float4 v0 = ldu(en, 0), v1 = ldu(en, 1);
          float p2, G;
          pair(r, w, c, p2, G);
          e += p2;
          wxx = fmaf(-v0.x, G, wxx);
          wyy = fmaf(-v0.y, G, wyy);
          wzz = fmaf(-v0.z, G, wzz);
          wxy = fmaf(-v0.w, G, wxy);
          wyz = fmaf(-v1.x, G, wyz);
          wzx = fmaf(-v1.y, G, wzx);
        });
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

  __device__ static double force_pass(const DevSet&, const PopCtl&, const Extra&, int, int, int, int)
  {
    return 0.0;
  }
};

void launch_nn2b_tabulate_f32(const PopCtl& ctl, const NN2BGrid& g, float* table, hipStream_t st)
{
  const dim3 block(256), grid((g.G + 255) / 256, ctl.P, g.codes);
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
  NN2BPop::Extra ex{g, table, (long)ctl.K * g.G * 3 * WAVE};
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
