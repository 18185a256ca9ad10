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
#include <type_traits>

#include "fastmath.h"
#include "kernels.h"
#include "population.h"

namespace gpuga {

// ------------------------------------------------------------ scalar pieces

// Per-parameter-set constants in the form the kernels use: exponent
// prefactors scaled by log2(e) for the exp2 instruction, 1/(R2-R1) and
// 0.5 pi/(R2-R1) for the taper.
struct TLane {
  float cR[2], kR[2], kR2[2], cA[2], kA[2], kA2[2], r0[2];
  float en, mhon, en1, mhon1, beta, beta2, h, r1, r2, inv_span, pif_half;
  float r1min; // population kernels: the smallest R1 of the wave's lanes
};

__host__ __device__ inline TLane tersoff_lane(const TersoffParams& p)
{
  TLane q;
  for (int k = 0; k < 2; ++k) {
    q.cR[k] = p.cR[k];
    q.kR[k] = p.kR[k];
    q.kR2[k] = p.kR[k] * LOG2E;
    q.cA[k] = p.cA[k];
    q.kA[k] = p.kA[k];
    q.kA2[k] = p.kA[k] * LOG2E;
    q.r0[k] = p.r0[k];
  }
  q.en = p.en;
  q.mhon = p.minus_half_over_n;
  q.en1 = p.en - 1.0f;
  q.mhon1 = p.minus_half_over_n - 1.0f;
  q.beta = p.beta;
  q.beta2 = 2.0f * p.beta;
  q.h = p.h;
  q.r1 = p.r1;
  q.r2 = p.r2;
  q.inv_span = 1.0f / (p.r2 - p.r1);
  q.pif_half = 0.5f * p.pi_factor;
  q.r1min = p.r1;
  return q;
}

// fc and dfc/dr (src/minimal_tersoff.cu:77-90)
__device__ __forceinline__ void tersoff_fc(const TLane& q, float d, float& fc, float& fcp)
{
  // 1 below r1, 0 from r2 on (a select), the taper between them under one
  // branch that the wave skips when no lane needs it
  const bool below = d < q.r1;
  fc = below ? 1.0f : 0.0f;
  fcp = 0.0f;
  if (!below && d < q.r2) {
    float sn, cs;
    sincospi01((d - q.r1) * q.inv_span, sn, cs);
    fc = 0.5f * cs + 0.5f;
    fcp = -q.pif_half * sn;
  }
}

// b = (1 + zeta^n)^(-1/2n) and db/dzeta = -1/2 (1 + zeta^n)^(-1/2n - 1)
// zeta^(n - 1), the latter as one exp2 of the summed logarithms (no
// reciprocal); b = 1, b' = 0 below zeta = 1e-16 (src/minimal_tersoff.cu:174-185)
__device__ __forceinline__ void tersoff_bond_order(const TLane& q, float zeta, float& b, float& bp)
{
  const float lz = flog2(zeta);
  const float bzn = fexp2(q.en * lz);
  const float l1 = flog2(1.0f + bzn);
  // the cut-off as selects of the exponents (exp2(0) = 1, exp2(-inf) = 0),
  // so that no branch forms around the exp2 instructions
  const bool tiny = zeta < 1.0e-16f;
  b = fexp2(tiny ? 0.0f : q.mhon * l1);
  bp = -fexp2(tiny ? -__builtin_inff() : fmaf(q.mhon1, l1, fmaf(q.en1, lz, -1.0f)));
}

// fR, fR', fA, fA' (src/minimal_tersoff.cu:62-75) of parameter set k
__device__ __forceinline__ void tersoff_pair(const TLane& q, int k, float d, float& fr, float& frp,
                                             float& fa, float& fap)
{
  const float x = d - q.r0[k];
  fr = q.cR[k] * fexp2(q.kR2[k] * x);
  frp = q.kR[k] * fr;
  fa = q.cA[k] * fexp2(q.kA2[k] * x);
  fap = q.kA[k] * fa;
}

// --------------------------------------- single-parameter-set kernels
// Thread per atom; scratch per bond: b, b' then f_ij.

__global__ void tersoff_plain_bond_order(DevSet s, TLane q, float* __restrict__ gb,
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

__global__ void tersoff_plain_partial_force(DevSet s, TLane q, const float* __restrict__ gb,
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
    const float dinv = frcp(r.w);
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
      const float g = q.beta * x * x, gp = q.beta2 * x;
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
  const TLane l = tersoff_lane(q);
  tersoff_plain_bond_order<<<grid, block, 0, st>>>(s, l, b, bp);
  tersoff_plain_partial_force<<<grid, block, 0, st>>>(s, l, b, bp, f12, pe);
  tersoff_plain_assemble<<<grid, block, 0, st>>>(s, f12, f, virial);
  GPUGA_HIP(hipGetLastError());
}

// ---------------------------------------------------------- population

// NMAX bounds the neighbour count of the training set so the per-bond values
// of an atom live in registers; TWO selects the two-species parameter sets.
template <int NMAX, bool TWO>
struct TersoffPop {
  static constexpr bool kForcePass = true;
  static constexpr int kFoldValues = 24;
  using Lane = TLane;
  using Extra = int;

  __device__ static Lane make(const PopCtl& ctl, int p)
  {
    constexpr int V = TWO ? 13 : 9;
    float x[V];
#pragma unroll
    for (int v = 0; v < V; ++v)
      x[v] = denormalise(ldu(ctl.pmin, v), ldu(ctl.pmax, v), ctl.genes[(size_t)p * V + v]);
    TLane q = tersoff_lane(tersoff_params(TWO ? 2 : 1, x));
    // the wave's smallest R1 (every lane calls make): an atom whose bonds
    // are all shorter has fc = 1, fc' = 0 on every lane
    float m = q.r1;
#pragma unroll
    for (int o = WAVE / 2; o >= 1; o >>= 1) m = fminf(m, __shfl_xor(m, o));
    q.r1min = m;
    return q;
  }

  // One atom for 64 individuals: the geometry is wave-uniform (scalar loads
  // from the atom's record), the parameters are per lane; all loop guards are
  // scalar branches (U = false: per-lane atoms of a packed item, ordinary
  // loads).
  // record stride in float4: header, bonds, triplets, reverse-bond indices,
  // inverse distances
  static constexpr int RS = 1 + NMAX + NMAX * NMAX + NMAX / 2;
  static constexpr int REV = 1 + NMAX + NMAX * NMAX; // reverse-bond indices
  static constexpr int INV = REV + NMAX / 4;          // inverse distances

  // record header: neighbour count, bond-type bits, longest bond
  struct Hdr {
    int nn, unl;
    float dmax;
  };

  template <bool FORCE, bool U>
  __device__ static void atom(const DevSet& s, const TLane& q, int i, float* fblk, int lane,
                              Acc& acc)
  {
    if constexpr (U) i = __builtin_amdgcn_readfirstlane(i);
    // header words as int loads (an int4 load through the constant address
    // space is issued as a vector load, waited for at every atom)
    const int* H = reinterpret_cast<const int*>(s.rec + (long)i * RS);
    atom_h<FORCE, U>(s, q, i, Hdr{ldx<U>(H, 0), TWO ? ldx<U>(H, 1) : 0, __int_as_float(ldx<U>(H, 3))},
                     s.rec + (long)i * RS, fblk, lane, acc);
  }

  // The next atom's record, touched by one vector load while the current
  // atom runs (a scalar load would be waited for at once: the kernel is
  // short of SGPRs, so the compiler moves such values to VGPR lanes as soon
  // as they arrive).  Lanes 0-2 read the header words (count, longest bond,
  // bond types), lanes 3-5 one word in each later 128-byte line of the
  // record, which brings those lines into L2 for the row loads.
  __device__ static int ahead_load(const DevSet& s, int i)
  {
    const char* base = reinterpret_cast<const char*>(s.rec + (long)i * RS);
    const int lane = __lane_id();
    const int off = lane == 1   ? 12
                    : lane == 2 ? 4
                    : lane == 3 ? min(128, RS * 16 - 4)
                    : lane == 4 ? min(256, RS * 16 - 4)
                    : lane == 5 ? RS * 16 - 4
                                : 0;
    return *reinterpret_cast<const int*>(base + off);
  }
  __device__ static Hdr header_of(int v)
  {
    return Hdr{__builtin_amdgcn_readlane(v, 0), TWO ? __builtin_amdgcn_readlane(v, 2) : 0,
               __int_as_float(__builtin_amdgcn_readlane(v, 1))};
  }

  // Row loads: scalar (U: wave-uniform record in global memory), from the
  // wave's LDS copy of the record (LR), or per lane.
  template <bool U, bool LR, class T>
  __device__ static T ldr(const T* p, long i)
  {
    if constexpr (U && !LR) return ldu(p, i);
    else return p[i];
  }

  // the atom with its header already loaded; R: its record (in LDS with LR)
  template <bool FORCE, bool U, bool LR = false>
  __device__ static void atom_h(const DevSet& s, const TLane& q, int i, const Hdr& h,
                                const float4* R, float* fblk, int lane, Acc& acc)
  {
    const int nn = h.nn, unl = h.unl;
    // an atom with the full neighbour bound (every atom of diamond Si) runs
    // a body without neighbour-count guards: no selects, no guarded loads;
    // when its longest bond is also shorter than every lane's R1 (the usual
    // case), fc = 1 and fc' = 0 exactly and the taper terms drop out
    // (four-bond records only: the larger bodies have no registers to spare
    // for a third copy)
    if constexpr (U) {
      if (nn == NMAX) {
        if (NMAX <= 4 && h.dmax < q.r1min)
          atom_body<FORCE, U, true, NMAX <= 4, LR>(s, q, i, R, nn, unl, fblk, lane, acc);
        else
          atom_body<FORCE, U, true, false, LR>(s, q, i, R, nn, unl, fblk, lane, acc);
        return;
      }
    }
    atom_body<FORCE, U, false, false, LR>(s, q, i, R, nn, unl, fblk, lane, acc);
  }

  // INNER: every bond inside every lane's R1.  Then fc = 1 and fc' = 0, and
  // the body leaves out what they multiply away: fc_v beta is beta,
  // B_v = A_v fc_v is A_v, the fc' terms of f_ij (and with them dsum) are
  // zero.  The results are the same bits as the general body's, which
  // multiplies by the exact 1 and adds the exact 0.
  template <bool FORCE, bool U, bool FULL, bool INNER, bool LR = false>
  __device__ static void atom_body(const DevSet& s, const TLane& q, int i, const float4* R,
                                   int nn, int unl, float* fblk, int lane, Acc& acc)
  {
    // Every load of a row is unconditional (records are zero-padded to the
    // bound), so the loads issue together and are waited for once, instead
    // of one load and wait per guarded use.  Uniform records (U) land in
    // SGPRs; per-lane records (packed items) in VGPRs, which the packed path
    // can afford: loaded at each use, inside the per-lane branches, every
    // load waited for on its own made a packed atom step cost five uniform
    // ones.
    constexpr bool PRE = NMAX <= 8; // preload rows (register budget allows)
    float d[NMAX], dinv[NMAX];
    if constexpr (PRE) {
#pragma unroll
      for (int u = 0; u < NMAX; ++u) {
        d[u] = ldr<U, LR>(R, 1 + u).w;
        dinv[u] = ldr<U, LR>(reinterpret_cast<const float*>(R + INV), u);
      }
    }
    auto dist = [&](int u) { return PRE ? d[u] : ldr<U, LR>(R, 1 + u).w; };
    auto dist_inv = [&](int u) {
      return PRE ? dinv[u] : ldr<U, LR>(reinterpret_cast<const float*>(R + INV), u);
    };
    float fc[NMAX], fcp[NMAX], fcb[NMAX], A[NMAX], B[NMAX], pf[NMAX];
#pragma unroll
    for (int u = 0; u < NMAX; ++u) {
      if (FULL || u < nn) {
        if constexpr (INNER) {
          fc[u] = 1.0f;
          fcp[u] = 0.0f;
          fcb[u] = q.beta;
        } else {
          tersoff_fc(q, dist(u), fc[u], fcp[u]);
          fcb[u] = fc[u] * q.beta;
        }
      }
    }
    // With four bonds, x = cos(theta_jk) - h and x^2 once per unordered pair:
    // the record holds the same cosine, bit for bit, at (j, k) and (k, j)
    // (the set-up computes it with commuting products in one order).
    constexpr bool SYM = NMAX <= 4;
    float xs[NMAX][NMAX], x2[NMAX][NMAX];
    if constexpr (SYM) {
#pragma unroll
      for (int u = 0; u < NMAX; ++u)
#pragma unroll
        for (int v = u + 1; v < NMAX; ++v)
          if (FULL || v < nn) {
            xs[u][v] = ldr<U, LR>(R, 1 + NMAX + u * NMAX + v).x - q.h;
            x2[u][v] = xs[u][v] * xs[u][v];
          }
    }
    auto xsym = [&](int u, int v) { return u < v ? xs[u][v] : xs[v][u]; };
    auto x2sym = [&](int u, int v) { return u < v ? x2[u][v] : x2[v][u]; };
    float e = 0.0f;
#pragma unroll
    for (int u = 0; u < NMAX; ++u) {
      if (FULL || u < nn) {
        float cs[NMAX];
        if constexpr (PRE && !SYM) {
#pragma unroll
          for (int v = 0; v < NMAX; ++v)
            if (v != u) cs[v] = ldr<U, LR>(R, 1 + NMAX + u * NMAX + v).x;
        }
        float zeta = 0.0f;
#pragma unroll
        for (int v = 0; v < NMAX; ++v) {
          if ((FULL || v < nn) && v != u) {
            if constexpr (SYM) {
              zeta = fmaf(fcb[v], x2sym(u, v), zeta);
            } else {
              const float x = (PRE ? cs[v] : ldr<U, LR>(R, 1 + NMAX + u * NMAX + v).x) - q.h;
              zeta = fmaf(fcb[v], x * x, zeta);
            }
          }
        }
        float b, bp;
        tersoff_bond_order(q, zeta, b, bp);
        // the parameter set as a branch on the (wave-uniform) bond type, so
        // the set is a compile-time index and the parameters stay in
        // registers (a run-time index puts them in scratch memory)
        float fr, frp, fa, fap;
        if (TWO && ((unl >> u) & 1)) tersoff_pair(q, 1, dist(u), fr, frp, fa, fap);
        else tersoff_pair(q, 0, dist(u), fr, frp, fa, fap);
        const float rep = fr - b * fa;
        const float di = dist_inv(u);
        A[u] = bp * fa;
        if constexpr (INNER) {
          e += rep;
          pf[u] = (frp - b * fap) * di;
          B[u] = A[u];
        } else {
          e = fmaf(fc[u], rep, e);
          pf[u] = (fcp[u] * rep + fc[u] * (frp - b * fap)) * di;
          fcp[u] *= di;
          B[u] = A[u] * fc[u];
        }
      }
    }
    float wxx = 0, wyy = 0, wzz = 0, wxy = 0, wyz = 0, wzx = 0;
#pragma unroll
    for (int u = 0; u < NMAX; ++u) {
      if (FULL || u < nn) {
        // three-body part of f_ij with beta factored out of both sums:
        //   c    = sum_k (h - cos)(A_j + A_k) fc_k dcos/dr   (times 2 beta)
        //   dsum = sum_k B_k (h - cos)^2                     (times beta)
        // where (A_j + A_k) fc_k = A_j fc_k + B_k
        float4 row[NMAX];
        if constexpr (PRE) {
#pragma unroll
          for (int v = 0; v < NMAX; ++v)
            if (v != u) row[v] = ldr<U, LR>(R, 1 + NMAX + u * NMAX + v);
        }
        float cx = 0.0f, cy = 0.0f, cz = 0.0f, dsum = 0.0f;
#pragma unroll
        for (int v = 0; v < NMAX; ++v) {
          if ((FULL || v < nn) && v != u) {
            const float4 t = PRE ? row[v] : ldr<U, LR>(R, 1 + NMAX + u * NMAX + v);
            const float x = SYM ? xsym(u, v) : t.x - q.h;
            const float w = (INNER ? A[u] + A[v] : fmaf(A[u], fc[v], B[v])) * x;
            cx = fmaf(w, t.y, cx);
            cy = fmaf(w, t.z, cy);
            cz = fmaf(w, t.w, cz);
            if constexpr (!INNER) {
              if constexpr (SYM) dsum = fmaf(B[v], x2sym(u, v), dsum);
              else dsum = fmaf(B[v] * x, x, dsum);
            }
          }
        }
        const float4 r = ldr<U, LR>(R, 1 + u);
        const float coef = INNER ? 0.5f * pf[u] : 0.5f * (pf[u] - fcp[u] * (q.beta * dsum));
        const float hc = INNER ? q.beta : fc[u] * q.beta;
        const float fx = coef * r.x - hc * cx;
        const float fy = coef * r.y - hc * cy;
        const float fz = coef * r.z - hc * cz;
        if (FORCE) {
          // f_ij goes to the slot of the reverse bond, where atom j reads it
          // back as f_ji; F_i's own part sum_j f_ij stays in registers
          const int rb = ldr<U, LR>(reinterpret_cast<const int*>(R + REV), u);
          float* o = fblk + (size_t)rb * 3 * WAVE + lane;
          cstore(o, fx);
          cstore(o + WAVE, fy);
          cstore(o + 2 * WAVE, fz);
          wxx += fx;
          wyy += fy;
          wzz += fz;
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
    if (FORCE) {
      // own part of F_i, in the per-atom slots after the B_force bond slots
      float* o = fblk + ((size_t)s.bond_force + i) * 3 * WAVE + lane;
      cstore(o, wxx);
      cstore(o + WAVE, wyy);
      cstore(o + 2 * WAVE, wzz);
    } else {
      acc.v[0] += (double)(0.5f * e);
      acc.v[1] += (double)wxx;
      acc.v[2] += (double)wyy;
      acc.v[3] += (double)wzz;
      acc.v[4] += (double)wxy;
      acc.v[5] += (double)wyz;
      acc.v[6] += (double)wzx;
    }
  }

  // Lone-wave form: a record spread over the wave, word l in lane l and
  // word 64 + l in lane l (two vector loads).
  static constexpr int RW = RS * 4; // words per record
  static_assert(NMAX > 4 || (RW >= WAVE && RW <= 2 * WAVE), "record layout of the lone-wave form");
  struct Spread {
    int lo, hi;
  };
  __device__ static Spread spread_load(const DevSet& s, int i)
  {
    const int* w = reinterpret_cast<const int*>(s.rec + (long)i * RS);
    const int l = __lane_id();
    return Spread{w[l], w[min(WAVE + l, RW - 1)]};
  }
  __device__ static void spread_store(float4* buf, const Spread& v)
  {
    int* w = reinterpret_cast<int*>(buf);
    const int l = __lane_id();
    w[l] = v.lo;
    if (WAVE + l < RW) w[WAVE + l] = v.hi;
  }
  __device__ static Hdr spread_header(const Spread& v)
  {
    return Hdr{__builtin_amdgcn_readlane(v.lo, 0), TWO ? __builtin_amdgcn_readlane(v.lo, 1) : 0,
               __int_as_float(__builtin_amdgcn_readlane(v.lo, 3))};
  }

  // HDR: load the next atom's header during the current atom (four-bond
  // records; the kernel instantiation without the packed path, whose
  // register budget has room for it).  LONE (launches of at most one wave
  // per SIMD, where a wave's time is the latency of its loads): the next
  // atom's whole record is loaded during the current atom, spread over the
  // wave, and written to the wave's LDS buffer at its end; the rows are read
  // from there (an LDS round trip per atom instead of a memory one).
  static constexpr bool kHeaderAhead = true;
  template <bool HDR = false, bool LONE = false>
  __device__ static void chunk(const DevSet& s, const TLane& q, const WorkItem& it,
                               const PopCtl& ctl, int, Acc& acc)
  {
    const int lane = __lane_id();
    const unsigned r = dispatch_round();
    if constexpr (LONE && NMAX <= 4) {
      // two records per wave of a 256-thread workgroup (the kernel's bound)
      __shared__ float4 buf[256 / WAVE][2][RS];
      const int w = wave_first((int)(threadIdx.x / WAVE));
      Spread cur = spread_load(s, it.a0);
      spread_store(buf[w][0], cur);
      auto run = [&](auto force) {
        constexpr bool F = decltype(force)::value;
        float* fblk = F ? ctl.bondf + (size_t)it.k * ctl.B_force * 3 * WAVE : nullptr;
        for (int i = it.a0; i < it.a1; ++i) {
          rotate_priority(r + (unsigned)i);
          const int b = (i - it.a0) & 1;
          const Hdr h = spread_header(cur);
          const Spread nxt = spread_load(s, min(i + 1, it.a1 - 1));
          atom_h<F, true, true>(s, q, i, h, buf[w][b], fblk, lane, acc);
          spread_store(buf[w][b ^ 1], nxt);
          cur = nxt;
        }
      };
      if (it.c < s.Nc_force) run(std::true_type{});
      else run(std::false_type{});
    } else if constexpr (NMAX > 4 || !HDR) {
      if (it.c < s.Nc_force) {
        float* fblk = ctl.bondf + (size_t)it.k * ctl.B_force * 3 * WAVE;
        for (int i = it.a0; i < it.a1; ++i) {
          rotate_priority(r + (unsigned)i);
          atom<true, true>(s, q, i, fblk, lane, acc);
        }
      } else {
        for (int i = it.a0; i < it.a1; ++i) {
          rotate_priority(r + (unsigned)i);
          atom<false, true>(s, q, i, nullptr, lane, acc);
        }
      }
    } else {
      // the next atom's header is loaded during this atom, so its row loads
      // issue at once (one memory latency per atom, not two)
      int hv = ahead_load(s, it.a0);
      if (it.c < s.Nc_force) {
        float* fblk = ctl.bondf + (size_t)it.k * ctl.B_force * 3 * WAVE;
        for (int i = it.a0; i < it.a1; ++i) {
          rotate_priority(r + (unsigned)i);
          const Hdr h = header_of(hv);
          hv = ahead_load(s, min(i + 1, it.a1 - 1));
          atom_h<true, true>(s, q, i, h, s.rec + (long)i * RS, fblk, lane, acc);
        }
      } else {
        for (int i = it.a0; i < it.a1; ++i) {
          rotate_priority(r + (unsigned)i);
          const Hdr h = header_of(hv);
          hv = ahead_load(s, min(i + 1, it.a1 - 1));
          atom_h<false, true>(s, q, i, h, s.rec + (long)i * RS, nullptr, lane, acc);
        }
      }
    }
  }

  // Packed item: lane group g takes the atoms a0 + g, a0 + g + G, ...; the
  // per-bond forces go to lane column col (the lane's individual).
  static constexpr bool kPacked = true;
  __device__ static void chunk_packed(const DevSet& s, const TLane& q, const WorkItem& it,
                                      const PopCtl& ctl, int, int col, int g, int G, Acc& acc)
  {
    if (it.c < s.Nc_force) {
      float* fblk = ctl.bondf + (size_t)it.k * ctl.B_force * 3 * WAVE;
      for (int i = it.a0 + g; i < it.a1; i += G) atom<true, false>(s, q, i, fblk, col, acc);
    } else {
      for (int i = it.a0 + g; i < it.a1; i += G) atom<false, false>(s, q, i, nullptr, col, acc);
    }
  }

  // F_i = sum_j (f_ij - f_ji) for the atoms [a0, a1) of a force configuration
  // from the per-bond forces every chunk of the configuration published, then
  // sum_i |F_i - F_ref,i|^2 (src/minimal_tersoff.cu:355-383,
  // src/fitness.cu:392-435).
  __device__ static double force_pass(const DevSet& s, const PopCtl& ctl, const Extra&, int a0,
                                      int a1, int k, int lane)
  {
    // per lane block: bond_force slots of received f_ji, then one slot per
    // force-configuration atom with its own sum_j f_ij (written in phase 0)
    const float* fblk = ctl.bondf + (size_t)k * ctl.B_force * 3 * WAVE + lane;
    const long N = s.N;
    double err = 0.0;
    // four atoms per step, every load unconditional, so all of them (60)
    // are in flight together: the loads go to memory, and their latency,
    // not their number, sets the time
    constexpr int STEP = 4;
    for (int i = a0; i < a1; i += STEP) {
      float d[STEP][3];
#pragma unroll
      for (int h = 0; h < STEP; ++h) {
        const int ii = min(i + h, a1 - 1);
        const int* H = reinterpret_cast<const int*>(s.rec + (long)ii * RS);
        const int nn = ldu(H, 0), b0 = ldu(H, 2);
        const float* own = fblk + ((size_t)s.bond_force + ii) * 3 * WAVE;
        float fx = cload(own), fy = cload(own + WAVE), fz = cload(own + 2 * WAVE);
#pragma unroll
        for (int u = 0; u < NMAX; ++u) {
          const bool on = u < nn;
          const float* r = fblk + (size_t)(on ? b0 + u : b0) * 3 * WAVE;
          const float rx = cload(r), ry = cload(r + WAVE), rz = cload(r + 2 * WAVE);
          fx -= on ? rx : 0.0f;
          fy -= on ? ry : 0.0f;
          fz -= on ? rz : 0.0f;
        }
        d[h][0] = fx - ldu(s.force_ref, ii);
        d[h][1] = fy - ldu(s.force_ref, ii + N);
        d[h][2] = fz - ldu(s.force_ref, ii + 2 * N);
      }
#pragma unroll
      for (int h = 0; h < STEP; ++h)
        if (i + h < a1)
          err += (double)d[h][0] * d[h][0] + (double)d[h][1] * d[h][1] +
                 (double)d[h][2] * d[h][2];
    }
    return err;
  }

  // The same for lane group g of a packed item (atoms a0 + g, a0 + g + G, ...
  // of lane column col).
  __device__ static double force_pass_packed(const DevSet& s, const PopCtl& ctl, int a0, int a1,
                                             int k, int col, int g, int G)
  {
    const float* fblk = ctl.bondf + (size_t)k * ctl.B_force * 3 * WAVE + col;
    const long N = s.N;
    double err = 0.0;
    for (int i = a0 + g; i < a1; i += G) {
      const int4 hd = reinterpret_cast<const int4*>(s.rec)[(long)i * RS];
      const float* own = fblk + ((size_t)s.bond_force + i) * 3 * WAVE;
      float fx = cload(own), fy = cload(own + WAVE), fz = cload(own + 2 * WAVE);
#pragma unroll
      for (int u = 0; u < NMAX; ++u) {
        if (u < hd.x) {
          const float* r = fblk + (size_t)(hd.z + u) * 3 * WAVE;
          fx -= cload(r);
          fy -= cload(r + WAVE);
          fz -= cload(r + 2 * WAVE);
        }
      }
      const float dx = fx - s.force_ref[i], dy = fy - s.force_ref[i + N],
                  dz = fz - s.force_ref[i + 2 * N];
      err += (double)dx * dx + (double)dy * dy + (double)dz * dz;
    }
    return err;
  }
};

// Sets with more than 16 neighbours within R2: the reference's bond loops
// (src/minimal_tersoff.cu:122-316) with run-time trip counts, the geometry
// read from the bond and triplet tables (wave-uniform loads) and the per-bond
// values of the 64 lanes (fc, fc'/d, b' fA, the pair part of f_ij) kept in
// LDS rows, lane-contiguous, one region per wave.
template <bool TWO>
struct TersoffLoop {
  static constexpr bool kForcePass = true;
  static constexpr int kFoldValues = 24;
  static constexpr bool kPacked = false;
  using Lane = TLane;
  using Extra = LoopTab;

  __device__ static Lane make(const PopCtl& ctl, int p) { return TersoffPop<4, TWO>::make(ctl, p); }

  template <bool FORCE>
  __device__ static void atom(const DevSet& s, const LoopTab& t, const TLane& q, int i,
                              float* fblk, int lane, float* L, Acc& acc)
  {
    const int nn = ldu(t.nn, i), b0 = ldu(t.b0, i), t0 = ldu(t.tri0, i);
    const int M = t.max_nn;
    float* FC = L + lane;                    // fc(d_v)
    float* FCPD = L + M * WAVE + lane;       // fc'(d_v) / d_v
    float* AB = L + 2 * M * WAVE + lane;     // b'_v fA_v
    float* PF = L + 3 * M * WAVE + lane;     // pair part of f_ij, over d
    for (int v = 0; v < nn; ++v) {
      const float d = ldu(t.bond, b0 + v).w;
      float fc, fcp;
      tersoff_fc(q, d, fc, fcp);
      FC[v * WAVE] = fc;
      FCPD[v * WAVE] = fcp * frcp(d);
    }
    float e = 0.0f;
    for (int u = 0; u < nn; ++u) {
      float zeta = 0.0f;
      for (int v = 0; v < nn; ++v) {
        if (v == u) continue;
        const float x = ldu(t.tri, t0 + u * nn + v).x - q.h;
        zeta = fmaf(FC[v * WAVE] * q.beta, x * x, zeta);
      }
      float b, bp;
      tersoff_bond_order(q, zeta, b, bp);
      const float d = ldu(t.bond, b0 + u).w;
      float fr, frp, fa, fap;
      if (TWO && (ldu(t.ptype, b0 + u) & 1)) tersoff_pair(q, 1, d, fr, frp, fa, fap);
      else tersoff_pair(q, 0, d, fr, frp, fa, fap);
      const float fc = FC[u * WAVE];
      const float rep = fr - b * fa;
      e = fmaf(fc, rep, e);
      // (fc' rep + fc (fR' - b fA')) / d, with fc' / d from the first loop
      PF[u * WAVE] = fmaf(FCPD[u * WAVE], rep, fc * (frp - b * fap) * frcp(d));
      AB[u * WAVE] = bp * fa;
    }
    float wxx = 0, wyy = 0, wzz = 0, wxy = 0, wyz = 0, wzx = 0;
    for (int u = 0; u < nn; ++u) {
      const float Au = AB[u * WAVE];
      float cx = 0.0f, cy = 0.0f, cz = 0.0f, dsum = 0.0f;
      for (int v = 0; v < nn; ++v) {
        if (v == u) continue;
        const float4 tr = ldu(t.tri, t0 + u * nn + v);
        const float x = tr.x - q.h;
        const float fcv = FC[v * WAVE];
        const float Bv = AB[v * WAVE] * fcv;
        const float w = fmaf(Au, fcv, Bv) * x;
        cx = fmaf(w, tr.y, cx);
        cy = fmaf(w, tr.z, cy);
        cz = fmaf(w, tr.w, cz);
        dsum = fmaf(Bv * x, x, dsum);
      }
      const float4 r = ldu(t.bond, b0 + u);
      const float coef = 0.5f * (PF[u * WAVE] - FCPD[u * WAVE] * (q.beta * dsum));
      const float hc = FC[u * WAVE] * q.beta;
      const float fx = coef * r.x - hc * cx;
      const float fy = coef * r.y - hc * cy;
      const float fz = coef * r.z - hc * cz;
      if (FORCE) {
        const int rb = ldu(t.rev, b0 + u);
        float* o = fblk + (size_t)rb * 3 * WAVE + lane;
        cstore(o, fx);
        cstore(o + WAVE, fy);
        cstore(o + 2 * WAVE, fz);
        wxx += fx;
        wyy += fy;
        wzz += fz;
      } else {
        wxx -= r.x * fx;
        wyy -= r.y * fy;
        wzz -= r.z * fz;
        wxy -= r.x * fy;
        wyz -= r.y * fz;
        wzx -= r.z * fx;
      }
    }
    if (FORCE) {
      float* o = fblk + ((size_t)s.bond_force + i) * 3 * WAVE + lane;
      cstore(o, wxx);
      cstore(o + WAVE, wyy);
      cstore(o + 2 * WAVE, wzz);
    } else {
      acc.v[0] += (double)(0.5f * e);
      acc.v[1] += (double)wxx;
      acc.v[2] += (double)wyy;
      acc.v[3] += (double)wzz;
      acc.v[4] += (double)wxy;
      acc.v[5] += (double)wyz;
      acc.v[6] += (double)wzx;
    }
  }

  static constexpr bool kHeaderAhead = false;
  __device__ static void chunk(const DevSet& s, const TLane& q, const WorkItem& it,
                               const PopCtl& ctl, const LoopTab& t, Acc& acc)
  {
    extern __shared__ float loop_lds[];
    const int lane = __lane_id();
    float* L = loop_lds + (size_t)(threadIdx.x / WAVE) * 4 * t.max_nn * WAVE;
    if (it.c < s.Nc_force) {
      float* fblk = ctl.bondf + (size_t)it.k * ctl.B_force * 3 * WAVE;
      for (int i = it.a0; i < it.a1; ++i) atom<true>(s, t, q, i, fblk, lane, L, acc);
    } else {
      for (int i = it.a0; i < it.a1; ++i) atom<false>(s, t, q, i, nullptr, lane, L, acc);
    }
  }

  // F_i = sum_j (f_ij - f_ji) over the list the phase-0 items used, then
  // sum_i |F_i - F_ref,i|^2 (src/minimal_tersoff.cu:355-383)
  __device__ static double force_pass(const DevSet& s, const PopCtl& ctl, const LoopTab& t,
                                      int a0, int a1, int k, int lane)
  {
    const float* fblk = ctl.bondf + (size_t)k * ctl.B_force * 3 * WAVE + lane;
    const long N = s.N;
    double err = 0.0;
    for (int i = a0; i < a1; ++i) {
      const int nn = ldu(t.nn, i), b0 = ldu(t.b0, i);
      const float* own = fblk + ((size_t)s.bond_force + i) * 3 * WAVE;
      float fx = cload(own), fy = cload(own + WAVE), fz = cload(own + 2 * WAVE);
      for (int u = 0; u < nn; ++u) {
        const float* r = fblk + (size_t)(b0 + u) * 3 * WAVE;
        fx -= cload(r);
        fy -= cload(r + WAVE);
        fz -= cload(r + 2 * WAVE);
      }
      const float dx = fx - ldu(s.force_ref, i), dy = fy - ldu(s.force_ref, i + N),
                  dz = fz - ldu(s.force_ref, i + 2 * N);
      err += (double)dx * dx + (double)dy * dy + (double)dz * dz;
    }
    return err;
  }
};

size_t tersoff_loop_lds_bytes(int max_nn, int block)
{
  return (size_t)(block / WAVE) * 4 * max_nn * WAVE * sizeof(float);
}

void launch_tersoff_loop_population(const PopCtl& ctl, const DevSet& s, const LoopTab& t,
                                    int grid, int block, hipStream_t st)
{
  const size_t lds = tersoff_loop_lds_bytes(t.max_nn, block);
  if (ctl.V == 13)
    population_kernel<TersoffLoop<true>, false><<<grid, block, lds, st>>>(ctl, s, t);
  else
    population_kernel<TersoffLoop<false>, false><<<grid, block, lds, st>>>(ctl, s, t);
  GPUGA_HIP(hipGetLastError());
}

int tersoff_loop_occupancy(bool two, int max_nn, int block)
{
  int n = 0;
  const void* f =
    two ? reinterpret_cast<const void*>(&population_kernel<TersoffLoop<true>, false>)
        : reinterpret_cast<const void*>(&population_kernel<TersoffLoop<false>, false>);
  GPUGA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block,
                                                         tersoff_loop_lds_bytes(max_nn, block)));
  return n;
}

template <int NMAX, bool TWO>
static void launch_t(const PopCtl& ctl, const DevSet& s, int grid, int block, bool pack,
                     bool lone, hipStream_t st)
{
  if (pack) {
    population_kernel<TersoffPop<NMAX, TWO>, true><<<grid, block, 0, st>>>(ctl, s, 0);
    return;
  }
  if constexpr (NMAX == 4) {
    if (lone) {
      population_kernel<TersoffPop<NMAX, TWO>, false, true><<<grid, block, 0, st>>>(ctl, s, 0);
      return;
    }
  }
  population_kernel<TersoffPop<NMAX, TWO>, false><<<grid, block, 0, st>>>(ctl, s, 0);
}

// cls: 0 for at most 4 neighbours per atom, 1 for 8, 2 for 16; pack: the
// packed-item instantiation; lone: the instantiation for launches of one
// wave per SIMD (four-bond records, no packed items)
void launch_tersoff_population(const PopCtl& ctl, const DevSet& s, int cls, int grid, int block,
                               bool pack, bool lone, hipStream_t st)
{
  const bool two = ctl.V == 13;
  switch (cls * 2 + (two ? 1 : 0)) {
  case 0: launch_t<4, false>(ctl, s, grid, block, pack, lone, st); break;
  case 1: launch_t<4, true>(ctl, s, grid, block, pack, lone, st); break;
  case 2: launch_t<8, false>(ctl, s, grid, block, pack, lone, st); break;
  case 3: launch_t<8, true>(ctl, s, grid, block, pack, lone, st); break;
  case 4: launch_t<16, false>(ctl, s, grid, block, pack, lone, st); break;
  default: launch_t<16, true>(ctl, s, grid, block, pack, lone, st); break;
  }
  GPUGA_HIP(hipGetLastError());
}

template <int NMAX, bool TWO>
static int occ_t(bool pack, int block)
{
  int n = 0;
  const void* f =
    pack ? reinterpret_cast<const void*>(&population_kernel<TersoffPop<NMAX, TWO>, true>)
         : reinterpret_cast<const void*>(&population_kernel<TersoffPop<NMAX, TWO>, false>);
  GPUGA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, 0));
  return n;
}

int tersoff_population_occupancy(int cls, bool two, bool pack, int block)
{
  switch (cls * 2 + (two ? 1 : 0)) {
  case 0: return occ_t<4, false>(pack, block);
  case 1: return occ_t<4, true>(pack, block);
  case 2: return occ_t<8, false>(pack, block);
  case 3: return occ_t<8, true>(pack, block);
  case 4: return occ_t<16, false>(pack, block);
  default: return occ_t<16, true>(pack, block);
  }
}

} // namespace gpuga
