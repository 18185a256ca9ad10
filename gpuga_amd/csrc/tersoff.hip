// Minimal Tersoff potential (one or two species) on CDNA.
//
// Two kernel families share the scalar pieces below:
//  * single-parameter-set kernels, one thread per atom, producing per-atom
//    force, virial and energy exactly as Potential::find_force does
//    (src/minimal_tersoff.cu:122-433); used by predict, find_force and the
//    per-individual ("plain") fitness path;
//  * the population kernel (population.h), one lane per individual, which
//    never materialises per-atom arrays for energy/virial configurations.
//
// Formulas: src/minimal_tersoff.cu:62-119 (fR, fA, fc, g), :147-185 (bond
// order), :250-307 (partial force f_ij), :355-392 (force and virial).
#include "kernels.h"
#include "population.h"

namespace gpuga {

// ------------------------------------------------------------ scalar pieces

__device__ inline void tersoff_fc(const TersoffParams& q, float d, float& fc, float& fcp)
{
  if (d < q.r1) {
    fc = 1.0f;
    fcp = 0.0f;
  } else if (d < q.r2) {
    float s, c;
    sincosf(q.pi_factor * (d - q.r1), &s, &c);
    fc = 0.5f * c + 0.5f;
    fcp = -s * q.pi_factor * 0.5f;
  } else {
    fc = 0.0f;
    fcp = 0.0f;
  }
}

// b = (1 + zeta^n)^(-1/2n), b' = db/dzeta; b = 1, b' = 0 below zeta = 1e-16
__device__ inline void tersoff_bond_order(const TersoffParams& q, float zeta, float& b, float& bp)
{
  if (zeta < 1.0e-16f) {
    b = 1.0f;
    bp = 0.0f;
  } else {
    const float bzn = exp2f(q.en * log2f(zeta));
    const float one = 1.0f + bzn;
    b = exp2f(q.minus_half_over_n * log2f(one));
    bp = -b * bzn * 0.5f / (one * zeta);
  }
}

__device__ inline void tersoff_pair(const TersoffParams& q, int unlike, float d, float& fr,
                                    float& frp, float& fa, float& fap)
{
  const float cR = unlike ? q.cR[1] : q.cR[0];
  const float kR = unlike ? q.kR[1] : q.kR[0];
  const float cA = unlike ? q.cA[1] : q.cA[0];
  const float kA = unlike ? q.kA[1] : q.kA[0];
  const float x = d - (unlike ? q.r0[1] : q.r0[0]);
  fr = cR * expf(kR * x);
  frp = kR * fr;
  fa = cA * expf(kA * x);
  fap = kA * fa;
}

// --------------------------------------- single-parameter-set kernels
// Thread per atom; scratch per bond: b, b' then f_ij.

__global__ void tersoff_plain_bond_order(DevSet s, TersoffParams q, float* __restrict__ gb,
                                         float* __restrict__ gbp)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  const int nn = s.nn[i], b0 = s.bond0[i];
  const float4* T = s.tri + s.tri0[i];
  for (int u = 0; u < nn; ++u) {
    float zeta = 0.0f;
    for (int v = 0; v < nn; ++v) {
      if (v == u) continue;
      float fc, fcp;
      tersoff_fc(q, s.bond[b0 + v].w, fc, fcp);
      const float x = T[u * nn + v].x - q.h;
      zeta += fc * (q.beta * x * x);
    }
    float b, bp;
    tersoff_bond_order(q, zeta, b, bp);
    gb[b0 + u] = b;
    gbp[b0 + u] = bp;
  }
}

__global__ void tersoff_plain_partial_force(DevSet s, TersoffParams q,
                                            const float* __restrict__ gb,
                                            const float* __restrict__ gbp,
                                            float* __restrict__ f12, float* __restrict__ pe)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  const int nn = s.nn[i], b0 = s.bond0[i];
  const float4* T = s.tri + s.tri0[i];
  float e = 0.0f;
  for (int u = 0; u < nn; ++u) {
    const float4 r = s.bond[b0 + u];
    const float dinv = 1.0f / r.w;
    float fc, fcp, fr, frp, fa, fap;
    tersoff_fc(q, r.w, fc, fcp);
    tersoff_pair(q, s.ptype[b0 + u], r.w, fr, frp, fa, fap);
    const float b = gb[b0 + u], bpu = gbp[b0 + u];
    const float au = bpu * fa;
    e += fc * (fr - b * fa) * 0.5f;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f, dsum = 0.0f;
    for (int v = 0; v < nn; ++v) {
      if (v == u) continue;
      float fcv, fcpv, frv, frpv, fav, fapv;
      tersoff_fc(q, s.bond[b0 + v].w, fcv, fcpv);
      tersoff_pair(q, s.ptype[b0 + v], s.bond[b0 + v].w, frv, frpv, fav, fapv);
      const float av = gbp[b0 + v] * fav;
      const float4 t = T[u * nn + v];
      const float x = t.x - q.h;
      const float g = q.beta * x * x, gp = 2.0f * q.beta * x;
      const float w = fcv * (au + av) * gp;
      cx += w * t.y;
      cy += w * t.z;
      cz += w * t.w;
      dsum += av * fcv * g;
    }
    const float pf = (fcp * (fr - b * fa) + fc * (frp - b * fap)) * dinv;
    const float coef = 0.5f * (pf - fcp * dinv * dsum);
    const int k = b0 + u;
    f12[3 * (size_t)k + 0] = coef * r.x - 0.5f * fc * cx;
    f12[3 * (size_t)k + 1] = coef * r.y - 0.5f * fc * cy;
    f12[3 * (size_t)k + 2] = coef * r.z - 0.5f * fc * cz;
  }
  pe[i] = e;
}

__global__ void tersoff_plain_assemble(DevSet s, const float* __restrict__ f12,
                                       float* __restrict__ f, float* __restrict__ virial)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  const int N = s.N, nn = s.nn[i], b0 = s.bond0[i];
  float fx = 0, fy = 0, fz = 0, wxx = 0, wyy = 0, wzz = 0, wxy = 0, wyz = 0, wzx = 0;
  for (int u = 0; u < nn; ++u) {
    const int k = b0 + u, kr = s.rev[k];
    const float4 r = s.bond[k];
    const float dx = f12[3 * (size_t)k + 0] - f12[3 * (size_t)kr + 0];
    const float dy = f12[3 * (size_t)k + 1] - f12[3 * (size_t)kr + 1];
    const float dz = f12[3 * (size_t)k + 2] - f12[3 * (size_t)kr + 2];
    fx += dx;
    fy += dy;
    fz += dz;
    wxx -= r.x * dx * 0.5f;
    wyy -= r.y * dy * 0.5f;
    wzz -= r.z * dz * 0.5f;
    wxy -= r.x * dy * 0.5f;
    wyz -= r.y * dz * 0.5f;
    wzx -= r.z * dx * 0.5f;
  }
  f[i] = fx;
  f[i + N] = fy;
  f[i + 2 * (size_t)N] = fz;
  virial[i] = wxx;
  virial[i + N] = wyy;
  virial[i + 2 * (size_t)N] = wzz;
  virial[i + 3 * (size_t)N] = wxy;
  virial[i + 4 * (size_t)N] = wyz;
  virial[i + 5 * (size_t)N] = wzx;
}

void launch_tersoff_plain(const DevSet& s, const TersoffParams& q, float* b, float* bp,
                          float* f12, float* f, float* virial, float* pe, hipStream_t st)
{
  const int block = 128, grid = (s.N + block - 1) / block;
  tersoff_plain_bond_order<<<grid, block, 0, st>>>(s, q, b, bp);
  tersoff_plain_partial_force<<<grid, block, 0, st>>>(s, q, b, bp, f12, pe);
  tersoff_plain_assemble<<<grid, block, 0, st>>>(s, f12, f, virial);
  GPUGA_HIP(hipGetLastError());
}

// ---------------------------------------------------------- population

struct TersoffPop {
  static constexpr bool kForcePass = true;
  using Lane = TersoffParams;
  using Extra = int;

  __device__ static Lane make(const PopCtl& ctl, int p)
  {
    float x[13];
#pragma unroll
    for (int v = 0; v < 13; ++v)
      x[v] = v < ctl.V ? denormalise(ctl.pmin[v], ctl.pmax[v], ctl.genes[(size_t)p * ctl.V + v])
                       : 0.0f;
    return tersoff_params(ctl.V == 13 ? 2 : 1, x);
  }

  // One atom for 64 individuals.  NMAX bounds the neighbour count so that the
  // per-bond values live in registers; all loop guards are wave-uniform.
  template <int NMAX, bool FORCE>
  __device__ static void atom(const DevSet& s, const TersoffParams& q, int i, float* fblk,
                              int lane, Acc& acc)
  {
    const int nn = s.nn[i], b0 = s.bond0[i];
    const float4* __restrict__ T = s.tri + s.tri0[i];
    float fc[NMAX], fcp[NMAX], A[NMAX], B[NMAX], pf[NMAX];
#pragma unroll
    for (int u = 0; u < NMAX; ++u) {
      if (u < nn) tersoff_fc(q, s.bond[b0 + u].w, fc[u], fcp[u]);
    }
    float e = 0.0f;
#pragma unroll
    for (int u = 0; u < NMAX; ++u) {
      if (u < nn) {
        float zeta = 0.0f;
#pragma unroll
        for (int v = 0; v < NMAX; ++v) {
          if (v < nn && v != u) {
            const float x = T[u * nn + v].x - q.h;
            zeta += fc[v] * (q.beta * x * x);
          }
        }
        float b, bp;
        tersoff_bond_order(q, zeta, b, bp);
        const float d = s.bond[b0 + u].w;
        float fr, frp, fa, fap;
        tersoff_pair(q, s.ptype[b0 + u], d, fr, frp, fa, fap);
        const float rep = fr - b * fa;
        e += fc[u] * rep;
        pf[u] = (fcp[u] * rep + fc[u] * (frp - b * fap)) / d;
        A[u] = bp * fa;
        B[u] = A[u] * fc[u];
      }
    }
    float wxx = 0, wyy = 0, wzz = 0, wxy = 0, wyz = 0, wzx = 0;
#pragma unroll
    for (int u = 0; u < NMAX; ++u) {
      if (u < nn) {
        float cx = 0.0f, cy = 0.0f, cz = 0.0f, dsum = 0.0f;
#pragma unroll
        for (int v = 0; v < NMAX; ++v) {
          if (v < nn && v != u) {
            const float4 t = T[u * nn + v];
            const float x = t.x - q.h;
            const float g = q.beta * x * x, gp = 2.0f * q.beta * x;
            const float w = fc[v] * (A[u] + A[v]) * gp;
            cx += w * t.y;
            cy += w * t.z;
            cz += w * t.w;
            dsum += B[v] * g;
          }
        }
        const float4 r = s.bond[b0 + u];
        const float coef = 0.5f * (pf[u] - fcp[u] / r.w * dsum);
        const float hc = 0.5f * fc[u];
        const float fx = coef * r.x - hc * cx;
        const float fy = coef * r.y - hc * cy;
        const float fz = coef * r.z - hc * cz;
        if (FORCE) {
          float* o = fblk + (size_t)(b0 + u) * 3 * WAVE + lane;
          o[0] = fx;
          o[WAVE] = fy;
          o[2 * WAVE] = fz;
        } else {
          // total virial of the configuration: sum_i W_i = -sum_ij r_ij (x) f_ij
          // because every bond's reverse carries -r_ij (exact FP32 minimum image)
          wxx -= r.x * fx;
          wyy -= r.y * fy;
          wzz -= r.z * fz;
          wxy -= r.x * fy;
          wyz -= r.y * fz;
          wzx -= r.z * fx;
        }
      }
    }
    if (!FORCE) {
      acc.v[0] += (double)(0.5f * e);
      acc.v[1] += (double)wxx;
      acc.v[2] += (double)wyy;
      acc.v[3] += (double)wzz;
      acc.v[4] += (double)wxy;
      acc.v[5] += (double)wyz;
      acc.v[6] += (double)wzx;
    }
  }

  template <int NMAX>
  __device__ static void run(const DevSet& s, const TersoffParams& q, const WorkItem& it,
                             const PopCtl& ctl, Acc& acc)
  {
    const int lane = __lane_id();
    if (it.c < s.Nc_force) {
      float* fblk = ctl.bondf + (size_t)it.k * ctl.B_force * 3 * WAVE;
      for (int i = it.a0; i < it.a1; ++i) atom<NMAX, true>(s, q, i, fblk, lane, acc);
    } else {
      for (int i = it.a0; i < it.a1; ++i) atom<NMAX, false>(s, q, i, nullptr, lane, acc);
    }
  }

  __device__ static void chunk(const DevSet& s, const TersoffParams& q, const WorkItem& it,
                               const PopCtl& ctl, int, Acc& acc)
  {
    switch (it.cls) {
    case 0: run<4>(s, q, it, ctl, acc); break;
    case 1: run<8>(s, q, it, ctl, acc); break;
    default: run<16>(s, q, it, ctl, acc); break;
    }
  }

  // F_i = sum_j (f_ij - f_ji) from the per-bond forces of every chunk of
  // configuration c, then sum_i |F_i - F_ref,i|^2 (src/minimal_tersoff.cu:355-383,
  // src/fitness.cu:392-435).
  __device__ static double force_pass(const DevSet& s, const PopCtl& ctl, int c, int k, int lane)
  {
    const float* fblk = ctl.bondf + (size_t)k * ctl.B_force * 3 * WAVE + lane;
    const int a0 = s.Na_sum[c], a1 = a0 + s.Na[c];
    const size_t N = s.N;
    double err = 0.0;
    for (int i = a0; i < a1; ++i) {
      const int nn = s.nn[i], b0 = s.bond0[i];
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
      for (int u = 0; u < nn; ++u) {
        const float* a = fblk + (size_t)(b0 + u) * 3 * WAVE;
        const float* r = fblk + (size_t)s.rev[b0 + u] * 3 * WAVE;
        fx += a[0] - r[0];
        fy += a[WAVE] - r[WAVE];
        fz += a[2 * WAVE] - r[2 * WAVE];
      }
      const float dx = fx - s.force_ref[i];
      const float dy = fy - s.force_ref[i + N];
      const float dz = fz - s.force_ref[i + 2 * N];
      err += (double)dx * dx + (double)dy * dy + (double)dz * dz;
    }
    return err;
  }
};

void launch_tersoff_population(const PopCtl& ctl, const DevSet& s, int grid, int block,
                               hipStream_t st)
{
  population_kernel<TersoffPop><<<grid, block, 0, st>>>(ctl, s, 0);
  GPUGA_HIP(hipGetLastError());
}

} // namespace gpuga

namespace gpuga {
int tersoff_population_occupancy(int block)
{
  int n = 0;
  GPUGA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, population_kernel<TersoffPop>, block, 0));
  return n;
}
} // namespace gpuga
