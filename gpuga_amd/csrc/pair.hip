// Pair potentials on CDNA: rigid ion (RI) and the NN2B neural pair potential.
//
//   RI    U, U'/r           src/ri.cu:77-110, pair loop :133-187
//   NN2B  MLP U, U'         src/nn2b.cu:49-78, cutoff :80-92, loop :115-171
//
// Single-parameter-set kernels keep the reference's thread-per-atom pair loop
// over the full neighbour list.  The RI population kernel reads, per bond,
// the parameter-independent factors erfc(a r)/r, its radial derivative,
// r^-6 and 1/r from a table built once per training set, so a lane only
// evaluates the parameter-dependent exponentials.
#include "kernels.h"
#include "population.h"

namespace gpuga {

// --------------------------------------------------------------------- RI

struct RIBond {
  float dinv6, ecoul, gcoul, dinv; // r^-6, erfc(a r)/r, -(d/dr)(erfc(a r)/r), 1/r
};

__device__ inline float sel3(const float* a, int t) { return t == 0 ? a[0] : (t == 1 ? a[1] : a[2]); }

// U(r) and U'(r)/r for a pair of type sum t (src/ri.cu:77-110).
__device__ inline void ri_pair(const RIParams& q, int t, float d, const float4& tab, float& p2,
                               float& f2)
{
  const float a = sel3(q.a, t), b = sel3(q.b, t), c = sel3(q.c, t), qq = sel3(q.qq, t);
  const float ex = expf(-d * b);
  float p = a * ex - c * tab.x;
  float g = 6.0f * c * tab.x * tab.w - a * ex * b;
  p += qq * (tab.y - q.v_rc - q.dv_rc * (d - q.cutoff));
  g -= qq * (tab.z + q.dv_rc);
  if (t == 1) {
    const float x = d - q.r0;
    const float fr = q.mR * expf(q.kR * x);
    const float fa = q.mA * expf(q.kA * x);
    p += fr - fa;
    g += q.kR * fr - q.kA * fa;
  }
  p2 = p;
  f2 = g * tab.w;
}

__global__ void ri_plain(DevSet s, RIParams q, const float4* __restrict__ tab,
                         float* __restrict__ f, float* __restrict__ virial, float* __restrict__ pe)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  const size_t N = s.N;
  const int nn = s.nn[i], b0 = s.bond0[i];
  float fx = 0, fy = 0, fz = 0, wxx = 0, wyy = 0, wzz = 0, wxy = 0, wyz = 0, wzx = 0, e = 0;
  for (int u = 0; u < nn; ++u) {
    const int k = b0 + u;
    const float4 r = s.bond[k];
    float p2, f2;
    ri_pair(q, s.ptype[k], r.w, tab[k], p2, f2);
    fx += r.x * f2;
    fy += r.y * f2;
    fz += r.z * f2;
    wxx -= r.x * r.x * f2 * 0.5f;
    wyy -= r.y * r.y * f2 * 0.5f;
    wzz -= r.z * r.z * f2 * 0.5f;
    wxy -= r.x * r.y * f2 * 0.5f;
    wyz -= r.y * r.z * f2 * 0.5f;
    wzx -= r.z * r.x * f2 * 0.5f;
    e += p2 * 0.5f;
  }
  f[i] = fx;
  f[i + N] = fy;
  f[i + 2 * N] = fz;
  virial[i] = wxx;
  virial[i + N] = wyy;
  virial[i + 2 * N] = wzz;
  virial[i + 3 * N] = wxy;
  virial[i + 4 * N] = wyz;
  virial[i + 5 * N] = wzx;
  pe[i] = e;
}

// Per-individual RI constants in the population form: exp(-d/rho) as
// exp2(-d b2) with b2 = log2(e)/rho, and the shifted Coulomb terms
// qq (erfc(a d)/d - v_rc - dv_rc (d - rc)) as qq erfc(a d)/d - k1 d - k0.
struct RILane {
  float a[3], b[3], b2[3], c[3], qq[3], k0[3], k1[3];
  float mR, kR, kR2, mA, kA, kA2, r0;
  int morse; // wave-uniform: some lane of the wave has a Morse term
  // wave-uniform: bit t set when some lane's pairs of type t have a term
  // evaluated per pair (A_t != 0, or the Morse term on unlike pairs); the
  // others add exact zeros and are left out
  int live;
};

__device__ inline RILane ri_lane(const RIParams& p)
{
  RILane q;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    q.a[t] = p.a[t];
    q.b[t] = p.b[t];
    q.b2[t] = p.b[t] * LOG2E;
    q.c[t] = p.c[t];
    q.qq[t] = p.qq[t];
    q.k0[t] = p.qq[t] * (p.v_rc - p.dv_rc * p.cutoff);
    q.k1[t] = p.qq[t] * p.dv_rc;
  }
  q.mR = p.mR;
  q.kR = p.kR;
  q.kR2 = p.kR * LOG2E;
  q.mA = p.mA;
  q.kA = p.kA;
  q.kA2 = p.kA * LOG2E;
  q.r0 = p.r0;
  q.morse = __any(p.mR != 0.0f || p.mA != 0.0f);
  q.live = (__any(p.a[0] != 0.0f) ? 1 : 0) | (__any(p.a[1] != 0.0f) || q.morse ? 2 : 0) |
           (__any(p.a[2] != 0.0f) ? 4 : 0);
  return q;
}

struct RIPop {
  static constexpr bool kForcePass = false;
  static constexpr int kFoldValues = 12;
  static constexpr bool kPacked = false;
  using Lane = RILane;
  using Extra = const float4*;

  __device__ static Lane make(const PopCtl& ctl, int p)
  {
    float x[15];
#pragma unroll
    for (int v = 0; v < 15; ++v)
      x[v] = denormalise(ldu(ctl.pmin, v), ldu(ctl.pmax, v), ctl.genes[(size_t)p * ctl.V + v]);
    return ri_lane(ri_params(x));
  }

  // The pairs of type T of one atom: the parts of F_i += u g (force
  // configurations) or of E += U, W -= u (x) u d g over the unordered pairs
  // (the others) that are not linear in the parameters: the Buckingham
  // exponential and, for unlike pairs, the Morse term (the linear parts come
  // from the set's sums, linear_force and linear_config).  The geometry is
  // wave-uniform (scalar loads); four pairs per step (their loads issue
  // together and are waited for once), then two, then one.  (float4 rows
  // per pair: one for a force configuration's pair, two for an energy
  // configuration's)
  template <int T, bool FORCE>
  __device__ static void pairs(const RILane& q, const float4* tab, int n, float* a)
  {
    constexpr int R = FORCE ? 1 : 2;
    int k = 0;
    for (; k + 4 <= n; k += 4) {
      float4 r[4][R];
#pragma unroll
      for (int h = 0; h < 4; ++h)
#pragma unroll
        for (int m = 0; m < R; ++m) r[h][m] = ldu(tab, R * (k + h) + m);
#pragma unroll
      for (int h = 0; h < 4; ++h) one<T, FORCE>(q, r[h], a);
    }
    for (; k + 2 <= n; k += 2) {
      float4 r[2][R];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int m = 0; m < R; ++m) r[h][m] = ldu(tab, R * (k + h) + m);
#pragma unroll
      for (int h = 0; h < 2; ++h) one<T, FORCE>(q, r[h], a);
    }
    if (k < n) {
      float4 r[R];
#pragma unroll
      for (int m = 0; m < R; ++m) r[m] = ldu(tab, R * k + m);
      one<T, FORCE>(q, r, a);
    }
  }

  // The non-linear terms of one pair (src/ri.cu:77-110): U = A e^(-d/rho)
  // (+ Morse) and g = U' = -A e^(-d/rho)/rho (+ Morse'); F += u g, or E += U,
  // W -= v g
  template <int T, bool FORCE>
  __device__ static __forceinline__ void one(const RILane& q, const float4 (&r)[FORCE ? 1 : 2],
                                             float* a)
  {
    const float d = FORCE ? r[0].w : r[0].x;
    const float aex = q.a[T] * fexp2(-d * q.b2[T]);
    float p = aex, g = -aex * q.b[T];
    if (T == 1 && q.morse) {
      const float x = d - q.r0;
      const float fr = q.mR * fexp2(q.kR2 * x);
      const float fa = q.mA * fexp2(q.kA2 * x);
      p += fr - fa;
      g += q.kR * fr - q.kA * fa;
    }
    if constexpr (FORCE) {
      // r[0] = (unit vector, d)
      a[0] = fmaf(r[0].x, g, a[0]);
      a[1] = fmaf(r[0].y, g, a[1]);
      a[2] = fmaf(r[0].z, g, a[2]);
    } else {
      // r[0] = (d, u_x u_x d, u_y u_y d, u_z u_z d), r[1] = (u_x u_y d,
      // u_y u_z d, u_z u_x d, 0)
      a[0] += p;
      a[1] = fmaf(-r[0].y, g, a[1]);
      a[2] = fmaf(-r[0].z, g, a[2]);
      a[3] = fmaf(-r[0].w, g, a[3]);
      a[4] = fmaf(-r[1].x, g, a[4]);
      a[5] = fmaf(-r[1].y, g, a[5]);
      a[6] = fmaf(-r[1].z, g, a[6]);
    }
  }

  // The linear part of F_i: per pair type t and component x,
  // C_t sum u_x 6 d^-7 - k1_t sum u_x - qq_t sum u_x h (dispersion, and the
  // shifted Coulomb g = -qq (h + dv_rc) with k1 = qq dv_rc).
  __device__ static void linear_force(const RILane& q, const float* w, float* a)
  {
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        const float* s = w + t * 9 + x * 3;
        a[x] = fmaf(q.c[t], ldu(s, 0), a[x]);
        a[x] = fmaf(-q.k1[t], ldu(s, 1), a[x]);
        a[x] = fmaf(-q.qq[t], ldu(s, 2), a[x]);
      }
  }

  // The linear part of a configuration's energy and virial (FP64, once per
  // configuration and lane block): per pair type t,
  //   E += qq_t sum erfc(a d)/d - C_t sum d^-6 - k1_t sum d - k0_t n
  //   W += -C_t sum v 6 d^-7 + k1_t sum v + qq_t sum v h
  __device__ static void linear_config(const RILane& q, const double* w, Acc& acc)
  {
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const double* s = w + t * kRiTypeConst;
      const double qq = q.qq[t], c = q.c[t], k1 = q.k1[t], k0 = q.k0[t];
      acc.v[0] += ((qq * ldu(s, 2) - c * ldu(s, 3)) - k1 * ldu(s, 1)) - k0 * ldu(s, 0);
#pragma unroll
      for (int m = 0; m < 6; ++m)
        acc.v[1 + m] += (k1 * ldu(s, 4 + m) - c * ldu(s, 10 + m)) + qq * ldu(s, 16 + m);
    }
  }

  // Energy/virial configurations visit each unordered pair once (the table
  // holds only j > i for their atoms: the reverse bond carries the same d and
  // -r); force configurations visit every bond so F_i is complete in the lane.
  static constexpr bool kHeaderAhead = false;
  __device__ static void chunk(const DevSet& s, const RILane& q, const WorkItem& it,
                               const PopCtl&, const float4*, Acc& acc)
  {
    const long N = s.N;
    if (it.c < s.Nc_force) {
      for (int i = it.a0; i < it.a1; ++i) {
        const int4 h = ldu(s.pair_hdr, i);
        const float4* tab = s.pair_tab + (long)h.x;
        float a[3] = {0.0f, 0.0f, 0.0f};
        linear_force(q, s.ri_fconst + (long)i * kRiAtomConst, a);
        if (q.live & 1) pairs<0, true>(q, tab, h.y, a);
        if (q.live & 2) pairs<1, true>(q, tab + (long)h.y, h.z, a);
        if (q.live & 4) pairs<2, true>(q, tab + (long)(h.y + h.z), h.w, a);
        const float dx = a[0] - ldu(s.force_ref, i), dy = a[1] - ldu(s.force_ref, i + N),
                    dz = a[2] - ldu(s.force_ref, i + 2 * N);
        acc.v[7] += (double)dx * dx + (double)dy * dy + (double)dz * dz;
      }
    } else {
      // the configuration's linear part goes with its first chunk
      if (it.chunk == 0) linear_config(q, s.ri_cconst + (long)it.c * 3 * kRiTypeConst, acc);
      for (int i = it.a0; i < it.a1; ++i) {
        const int4 h = ldu(s.pair_hdr, i);
        const float4* tab = s.pair_tab + (long)h.x;
        float a[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (q.live & 1) pairs<0, false>(q, tab, h.y, a);
        if (q.live & 2) pairs<1, false>(q, tab + 2 * (long)h.y, h.z, a);
        if (q.live & 4) pairs<2, false>(q, tab + 2 * (long)(h.y + h.z), h.w, a);
#pragma unroll
        for (int u = 0; u < 7; ++u) acc.v[u] += (double)a[u];
      }
    }
  }

  __device__ static double force_pass(const DevSet&, const PopCtl&, const Extra&, int, int, int,
                                      int)
  {
    return 0.0;
  }
};

void launch_ri_plain(const DevSet& s, const RIParams& q, const float4* tab, float* f,
                         float* virial, float* pe, hipStream_t st)
{
  const int block = 128, grid = (s.N + block - 1) / block;
  ri_plain<<<grid, block, 0, st>>>(s, q, tab, f, virial, pe);
  GPUGA_HIP(hipGetLastError());
}

void launch_ri_population(const PopCtl& ctl, const DevSet& s, const float4* ri_tab, int grid,
                          int block, hipStream_t st)
{
  population_kernel<RIPop><<<grid, block, 0, st>>>(ctl, s, ri_tab);
  GPUGA_HIP(hipGetLastError());
}

int ri_population_occupancy(int block)
{
  int n = 0;
  GPUGA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, population_kernel<RIPop>, block, 0));
  return n;
}

// -------------------------------------------------------------------- NN2B

constexpr int NN2B_MAX_PLAIN = 64;

// One thread per atom, weights staged in LDS; the MLP is evaluated per bond as
// the reference does (src/nn2b.cu:49-78), for any n <= 64.  species: the
// first layer also takes the pair's species code t_i + t_j through the n
// weights after the reference's parameters.
__global__ void nn2b_plain(DevSet s, int n, int species, const float* __restrict__ w,
                           float* __restrict__ f, float* __restrict__ virial,
                           float* __restrict__ pe)
{
  extern __shared__ float sw[];
  const int V = n * (n + 4) + 1 + (species ? n : 0);
  for (int v = threadIdx.x; v < V; v += blockDim.x) sw[v] = w[v];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.N) return;
  const float* w0 = sw;
  const float* b0 = sw + n;
  const float* w1 = sw + 2 * n;
  const float* b1 = sw + n * (n + 2);
  const float* w2 = sw + n * (n + 3);
  const float bias2 = sw[n * (n + 4)];
  const float* ws = sw + n * (n + 4) + 1;
  const size_t N = s.N;
  const int nn = s.nn[i], k0 = s.bond0[i];
  float fx = 0, fy = 0, fz = 0, wxx = 0, wyy = 0, wzz = 0, wxy = 0, wyz = 0, wzx = 0, e = 0;
  float x1[NN2B_MAX_PLAIN], y1[NN2B_MAX_PLAIN];
  for (int u = 0; u < nn; ++u) {
    const float4 r = s.bond[k0 + u];
    const float d = r.w;
    const float code = species ? (float)s.ptype[k0 + u] : 0.0f;
    for (int m = 0; m < n; ++m) {
      // the species-pair input as a shift of the bias
      const float bm = species ? b0[m] - ws[m] * code : b0[m];
      x1[m] = tanhf(w0[m] * d - bm);
      y1[m] = (1.0f - x1[m] * x1[m]) * w0[m];
    }
    float p = 0.0f, g = 0.0f;
    for (int a = 0; a < n; ++a) {
      float z = 0.0f, y = 0.0f;
      for (int m = 0; m < n; ++m) {
        z += w1[a * n + m] * x1[m];
        y += w1[a * n + m] * y1[m];
      }
      const float x2 = tanhf(z - b1[a]);
      y *= 1.0f - x2 * x2;
      p += w2[a] * x2;
      g += w2[a] * y;
    }
    p -= bias2;
    float fc, fcp;
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
    const float p2 = p * fc;
    const float f2 = (g * fc + p2 * fcp) / d; // src/nn2b.cu:147-148
    fx += r.x * f2;
    fy += r.y * f2;
    fz += r.z * f2;
    wxx -= r.x * r.x * f2 * 0.5f;
    wyy -= r.y * r.y * f2 * 0.5f;
    wzz -= r.z * r.z * f2 * 0.5f;
    wxy -= r.x * r.y * f2 * 0.5f;
    wyz -= r.y * r.z * f2 * 0.5f;
    wzx -= r.z * r.x * f2 * 0.5f;
    e += p2 * 0.5f;
  }
  f[i] = fx;
  f[i + N] = fy;
  f[i + 2 * N] = fz;
  virial[i] = wxx;
  virial[i + N] = wyy;
  virial[i + 2 * N] = wzz;
  virial[i + 3 * N] = wxy;
  virial[i + 4 * N] = wyz;
  virial[i + 5 * N] = wzx;
  pe[i] = e;
}

void launch_nn2b_plain(const DevSet& s, int n, bool species, const float* w, float* f,
                       float* virial, float* pe, hipStream_t st)
{
  if (n > NN2B_MAX_PLAIN) fail(strfmt("NN2B supports at most %d neurons", NN2B_MAX_PLAIN));
  const int block = 128, grid = (s.N + block - 1) / block;
  const size_t lds = sizeof(float) * (n * (n + 4) + 1 + (species ? n : 0));
  nn2b_plain<<<grid, block, lds, st>>>(s, n, species ? 1 : 0, w, f, virial, pe);
  GPUGA_HIP(hipGetLastError());
}

} // namespace gpuga
