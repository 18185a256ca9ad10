// Population kernel framework: one persistent launch evaluates every
// (individual, configuration) pair of a generation and reduces it to one
// fitness per individual.
//
// Mapping.  A wavefront owns 64 individuals ("lane block" k: individuals
// 64k .. 64k+63, one per lane) and walks the atoms of one configuration chunk.
// The training-set geometry an atom needs is identical for all 64 lanes, so it
// is read with wave-uniform (scalar) loads; only the potential parameters
// differ per lane.  Per-configuration energy and virial are accumulated per
// lane in FP64, so no cross-lane reduction is ever needed.
//
// Schedule.  Work items (configuration chunk x lane block) are sorted by cost,
// largest first, and pulled from a global counter by the resident waves.  The
// last chunk of a (configuration, lane block) to finish folds the chunk
// partials in chunk order and writes that configuration's error terms; the
// last configuration of a lane block to finish sums the error terms over the
// configurations in index order and writes the fitness.  Every reduction has a
// fixed order, so results do not depend on scheduling or on what else runs on
// the device.  All counters are returned to zero by their last user, so the
// launch can be replayed from a captured graph without clearing memory.
//
// Replaces the per-individual host loop of Fitness::compute
// (src/fitness.cu:316-333) and its reductions (src/fitness.cu:392-555).
#pragma once

#include "device.h"
#include "fastmath.h"

namespace gpuga {

struct WorkItem {
  int c;      // configuration
  int k;      // lane block
  int chunk;  // chunk index inside the configuration
  int a0, a1; // global atom range of the chunk
  int cls;    // neighbour-count class of the chunk
  int pad0, pad1;
};

struct PopCtl {
  const WorkItem* items;
  int n_items;
  int K;                  // lane blocks
  int P, V;
  const int* cfg_nchunks; // [Nc]
  const int* cfg_slot0;   // [Nc] first partial slot of configuration c
  const float* genes;     // [P][V] in [0,1]
  const float* pmin;      // [V]
  const float* pmax;      // [V]
  double* part;           // [slot][K][8][64]
  float* bondf;           // [K][B_force][3][64] per-bond forces (three-body only)
  int B_force;
  double* err;            // [K][3][Nc][64]
  unsigned* cnt;          // [Nc][K]
  unsigned* blk_done;     // [K]
  unsigned* fetch;        // [1]
  float* fitness;         // [P]
  double* terms;          // [P][3]
  FitnessNorm norm;
};

// Partial sums carried by one lane over a chunk: E, W[6], sum |F - F_ref|^2.
struct Acc {
  double v[8];
};

__device__ inline unsigned wave_atomic_add(unsigned* p, unsigned x)
{
  unsigned old = 0;
  if (__lane_id() == 0) old = atomicAdd(p, x);
  return (unsigned)wave_first((int)old);
}

// Lane block k, all configurations finished: sum the error terms in
// configuration order and write fitness and the three relative RMS errors.
__device__ inline void finish_block(const PopCtl& ctl, int Nc, int k, int lane)
{
  const double* e0 = ctl.err + ((size_t)(k * 3 + 0) * Nc) * WAVE + lane;
  const double* e1 = ctl.err + ((size_t)(k * 3 + 1) * Nc) * WAVE + lane;
  const double* e2 = ctl.err + ((size_t)(k * 3 + 2) * Nc) * WAVE + lane;
  double s0[4] = {0, 0, 0, 0}, s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
  int c = 0;
  for (; c + 4 <= Nc; c += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      s0[u] += e0[(size_t)(c + u) * WAVE];
      s1[u] += e1[(size_t)(c + u) * WAVE];
      s2[u] += e2[(size_t)(c + u) * WAVE];
    }
  }
  for (; c < Nc; ++c) {
    s0[0] += e0[(size_t)c * WAVE];
    s1[0] += e1[(size_t)c * WAVE];
    s2[0] += e2[(size_t)c * WAVE];
  }
  const double se = (s0[0] + s0[1]) + (s0[2] + s0[3]);
  const double sv = (s1[0] + s1[1]) + (s1[2] + s1[3]);
  const double sf = (s2[0] + s2[1]) + (s2[2] + s2[3]);
  const double tE = sqrt(se * ctl.norm.inv_pss);
  const double tW = sqrt(sv * ctl.norm.inv_vss);
  const double tF = sqrt(sf * ctl.norm.inv_fss);
  const int p = k * WAVE + lane;
  if (p < ctl.P) {
    ctl.fitness[p] =
      (float)(ctl.norm.w_energy * tE + ctl.norm.w_stress * tW + ctl.norm.w_force * tF);
    ctl.terms[3 * (size_t)p + 0] = tE;
    ctl.terms[3 * (size_t)p + 1] = tW;
    ctl.terms[3 * (size_t)p + 2] = tF;
  }
}

// Pot must provide:
//   struct Lane;                                   per-lane constants
//   static Lane make(ctl, p);                      denormalised genes of p
//   typedef Extra;                                 potential-specific tables
//   static void chunk(set, q, item, ctl, ex, Acc&) process atoms [a0, a1)
//   static void force_pass(set, ctl, c, k, lane)   -> sum |F-F_ref|^2 over
//        configuration c; only called when Pot::kForcePass is true
template <class Pot>
__global__ void __launch_bounds__(256) population_kernel(PopCtl ctl, DevSet set,
                                                         typename Pot::Extra ex)
{
  const int lane = __lane_id();
  const int waves_per_block = blockDim.x / WAVE;
  const unsigned total_waves = gridDim.x * waves_per_block;
  const int Nc = set.Nc;
  for (;;) {
    const unsigned it = wave_atomic_add(ctl.fetch, 1u);
    if (it >= (unsigned)ctl.n_items) {
      // The last fetch of the launch returns the counter to zero.
      if (it == (unsigned)ctl.n_items + total_waves - 1u && lane == 0) *ctl.fetch = 0u;
      break;
    }
    const WorkItem item = ldu(ctl.items, it);
    const int k = item.k;
    // individual of this lane; lanes past P repeat the last one and write nothing
    const int p = min(k * WAVE + lane, ctl.P - 1);
    const typename Pot::Lane q = Pot::make(ctl, p);

    Acc acc;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc.v[u] = 0.0;
    Pot::chunk(set, q, item, ctl, ex, acc);

    const int c = item.c;
    const int nch = ldu(ctl.cfg_nchunks, c);
    const bool is_force = c < set.Nc_force;
    if (nch > 1) {
      double* slot = ctl.part + ((size_t)(ldu(ctl.cfg_slot0, c) + item.chunk) * ctl.K + k) * 8 * WAVE;
#pragma unroll
      for (int u = 0; u < 8; ++u) slot[u * WAVE + lane] = acc.v[u];
      __threadfence();
      const unsigned done = wave_atomic_add(&ctl.cnt[(size_t)c * ctl.K + k], 1u);
      if (done != (unsigned)nch - 1u) continue;
      __threadfence();
      if (lane == 0) ctl.cnt[(size_t)c * ctl.K + k] = 0u;
      // fold the chunk partials in chunk order
#pragma unroll
      for (int u = 0; u < 8; ++u) acc.v[u] = 0.0;
      for (int ch = 0; ch < nch; ++ch) {
        const double* s = ctl.part + ((size_t)(ldu(ctl.cfg_slot0, c) + ch) * ctl.K + k) * 8 * WAVE;
#pragma unroll
        for (int u = 0; u < 8; ++u) acc.v[u] += s[u * WAVE + lane];
      }
    }
    double eE = 0.0, eW = 0.0, eF = 0.0;
    if (is_force) {
      eF = Pot::kForcePass ? Pot::force_pass(set, ctl, c, k, lane) : acc.v[7];
    } else {
      const double dE = acc.v[0] - (double)ldu(set.pe_ref, c);
      eE = dE * dE;
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        const double dW = acc.v[1 + u] - (double)ldu(set.virial_ref, c + (long)Nc * u);
        eW += dW * dW;
      }
    }
    ctl.err[((size_t)(k * 3 + 0) * Nc + c) * WAVE + lane] = eE;
    ctl.err[((size_t)(k * 3 + 1) * Nc + c) * WAVE + lane] = eW;
    ctl.err[((size_t)(k * 3 + 2) * Nc + c) * WAVE + lane] = eF;
    __threadfence();
    const unsigned cdone = wave_atomic_add(&ctl.blk_done[k], 1u);
    if (cdone != (unsigned)Nc - 1u) continue;
    __threadfence();
    if (lane == 0) ctl.blk_done[k] = 0u;
    finish_block(ctl, Nc, k, lane);
  }
}

} // namespace gpuga
