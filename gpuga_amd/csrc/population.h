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
// Schedule.  The resident waves pull work items from queues (a few per XCD;
// each configuration belongs to one queue, so its atom records and per-bond
// forces mostly stay in one L2), each queue with its own counter: a wave
// serves a queue of its own XCD and, once that is empty, the others.
// Phase-0 items (configuration chunk x lane block, largest chunk first, the
// lane blocks of a chunk together) evaluate the potential.  For three-body potentials the force on an atom needs the
// partial forces of its neighbours, so phase-0 items of force configurations
// only publish per-bond forces, and phase-1 items assemble F_i - F_ref,i for
// their chunk once all chunks of their configuration have published (a
// configuration of one chunk is assembled by the wave that evaluated it,
// right after, with no phase-1 item).  A
// phase-1 item sits in its configuration's queue after every phase-0 item of
// that configuration (about one item per resident wave of the queue later,
// so that they have finished by then).  Items of a queue are handed out in
// queue order to running waves and phase-0 work never waits, so every wait
// ends; it is bounded anyway, and a timeout is reported through a status
// word instead of hanging.
//
// Packed lane block.  When the last lane block holds at most 32 individuals
// (P mod 64), its items split the wave into G = 64 / W groups of W lanes
// (W the power of two >= that count): lane l evaluates individual 64k + l % W
// on the atoms a0 + l / W, a0 + l / W + G, ... of the chunk, with per-lane
// geometry loads, and the G group partials are summed by a fixed butterfly
// before the chunk result is stored.  Afterwards lane column l holds the
// result of individual 64k + l % W, so every later step is unchanged (the
// columns past the block's individuals are never read).
//
// Reductions.  The last chunk of a (configuration, lane block) folds the chunk
// partials in chunk order and writes the configuration's error terms; the last
// configuration of a group (about sqrt(Nc) of them) sums the group in index
// order; the last group of a lane block sums the groups and writes the
// fitness (with one group, its last configuration does).  Every reduction
// has a fixed order, so results do not depend on scheduling or on what else
// runs on the device.  All counters are returned to zero by their last user,
// so the launch can be replayed from a captured graph without clearing
// memory.
//
// Replaces the per-individual host loop of Fitness::compute
// (src/fitness.cu:316-333) and its reductions (src/fitness.cu:392-555).
#pragma once

#include "device.h"
#include "fastmath.h"

namespace gpuga {

// Configurations per reduction group: about sqrt(Nc), a power of two in
// [8, 64], so that the group fold and the lane-block fold over the groups
// are both a few memory round trips.  (One group for a set of at most 64
// configurations, with the lane-block fold left out, was slower on a set of
// 60: the group's last configuration then loads all 60 rows.)
inline int config_group(int Nc)
{
  int g = 8;
  while (g < 64 && g * g < Nc) g <<= 1;
  return g;
}
constexpr int kTraceFields = 6; // words per item of the optional timeline
constexpr int kQueuesPerXcd = 4; // work queues per XCD (default)
constexpr int kXcds = 8;         // XCDs of an MI355X
constexpr int kQueues = kQueuesPerXcd * kXcds;
// distance of two queue counters in 32-bit words (4 KiB + 256 B: distinct
// lines and, with the address interleave, distinct memory channels)
constexpr int kFetchStride = 1088;

struct WorkItem {
  int c;      // configuration
  int k;      // lane block
  int chunk;  // chunk index inside the configuration
  int a0, a1; // global atom range of the chunk
  int cls;    // neighbour-count class of the chunk
  int phase;  // 0: evaluate, 1: assemble forces of a three-body force configuration
  int groups; // 1: lane = individual; G > 1: packed lane block (see below)
};

struct PopCtl {
  const WorkItem* items;  // the queues, one after another
  int n_items;
  int nq;                 // queues
  const int* qstart;      // [nq + 1] first item of each queue
  const int* home;        // [n_home] first queue of wave slot w (else by XCD)
  int n_home;
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
  double* gpart;          // [K][groups][3][64]
  int cfg_group;          // configurations per reduction group
  unsigned* cnt;          // [Nc][K]  phase-0 chunks done
  unsigned* cnt2;         // [Nc][K]  phase-1 chunks done
  unsigned* grp_done;     // [K][groups]
  unsigned* blk_done;     // [K]
  unsigned* fetch;        // [nq] items handed out per queue, kFetchStride apart
  unsigned* done;         // [1] waves finished
  unsigned long long* closed; // bit q: every item of queue q has been handed out
  unsigned* status;       // [1] non-zero: a dependency wait timed out
  int spin_limit;         // polls of a dependency wait before it gives up
                          // (about 1 us each); < 0: give up at once
  float* fitness;         // [P]
  double* terms;          // [P][3]
  FitnessNorm norm;
  // optional item timeline: [n_items][kTraceFields] = wave slot, start,
  // end, the wave's entry (first item, else 0), compute begin, compute end
  // in wall-clock ticks (constant rate); null when not recording
  unsigned long long* trace;
};

// Partial sums carried by one lane over a chunk: E, W[6], sum |F - F_ref|^2.
struct Acc {
  double v[8];
};

// Cross-wave data (chunk partials, error terms, per-bond forces) is written
// and read with device-scope relaxed atomics: the accesses are coherent
// across the XCDs (they do not rest in an XCD's L2), so no cache-wide write
// back or invalidate is ever needed.  publish() waits until every such store
// of the wave has completed before the counter update that announces it;
// observe() keeps the compiler from moving the reads before the counter
// check that guards them (the hardware issues them after it).
template <class T>
__device__ __forceinline__ void cstore(T* p, T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ __forceinline__ T cload(const T* p)
{
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void publish() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void observe() { asm volatile("" ::: "memory"); }

__device__ inline unsigned wave_atomic_add(unsigned* p, unsigned x)
{
  unsigned old = 0;
  if (__lane_id() == 0) old = atomicAdd(p, x);
  return (unsigned)wave_first((int)old);
}

// Default poll count of a dependency wait (about a second).
constexpr int kSpinLimit = 1 << 23;

// Wait until *p >= target, polling at most `limit` more times after the
// first look; false on timeout.  limit < 0 reports a timeout without
// looking (the host sets it only to exercise the failure path).
__device__ inline bool wave_wait_count(const unsigned* p, unsigned target, int limit)
{
  if (limit < 0) return false;
  for (int spin = 0;; ++spin) {
    unsigned v = 0;
    if (__lane_id() == 0) v = cload(p);
    v = (unsigned)wave_first((int)v);
    if (v >= target) {
      observe();
      return true;
    }
    if (spin >= limit) return false;
    __builtin_amdgcn_s_sleep(8);
  }
}

// Sum n rows (stride `stride` doubles) of NA arrays in a fixed order.  The
// rows are read from memory (coherent loads), so the time is the number of
// round trips: VALS values (VALS / NA rows) are loaded before any is added,
// then the next VALS, into 4 partial sums per array.  (VALS sets the
// registers this needs: two per value.)
template <int NA, int VALS>
__device__ inline void sum_rows(const double* const (&a)[NA], long stride, int n, double out[NA])
{
  constexpr int B = VALS / NA;
  double s[NA][4];
#pragma unroll
  for (int x = 0; x < NA; ++x)
#pragma unroll
    for (int u = 0; u < 4; ++u) s[x][u] = 0.0;
  for (int c = 0; c < n; c += B) {
    double v[NA][B];
#pragma unroll
    for (int u = 0; u < B; ++u)
#pragma unroll
      for (int x = 0; x < NA; ++x) v[x][u] = c + u < n ? cload(a[x] + (c + u) * stride) : 0.0;
#pragma unroll
    for (int u = 0; u < B; ++u)
#pragma unroll
      for (int x = 0; x < NA; ++x) s[x][u & 3] += v[x][u];
  }
#pragma unroll
  for (int x = 0; x < NA; ++x) out[x] = (s[x][0] + s[x][1]) + (s[x][2] + s[x][3]);
}

// Fitness and terms of the lane block's individuals from the sums of the
// three error terms over all configurations.
__device__ inline void write_fitness(const PopCtl& ctl, int k, int lane, const double (&s)[3])
{
  const double tE = sqrt(s[0] * ctl.norm.inv_pss);
  const double tW = sqrt(s[1] * ctl.norm.inv_vss);
  const double tF = sqrt(s[2] * ctl.norm.inv_fss);
  const int p = k * WAVE + lane;
  if (p < ctl.P) {
    ctl.fitness[p] =
      (float)(ctl.norm.w_energy * tE + ctl.norm.w_stress * tW + ctl.norm.w_force * tF);
    ctl.terms[3 * (long)p + 0] = tE;
    ctl.terms[3 * (long)p + 1] = tW;
    ctl.terms[3 * (long)p + 2] = tF;
  }
}

// Configuration c of lane block k has its error terms (force configurations
// the force term eF only, the others energy and virial eE, eW): propagate
// through the group and lane-block reductions, writing the fitness when the
// lane block is complete (then it returns true).  Only the terms a
// configuration has are stored and read back.
template <int VALS>
__device__ inline bool config_done(const PopCtl& ctl, int Nc, int Nc_force, int c, int k,
                                   int lane, double eE, double eW, double eF)
{
  const long E3 = (long)Nc * WAVE;
  double* e = ctl.err + (long)k * 3 * E3 + (long)c * WAVE + lane;
  if (c < Nc_force) {
    cstore(e + 2 * E3, eF);
  } else {
    cstore(e, eE);
    cstore(e + E3, eW);
  }
  const int cg = ctl.cfg_group;
  const int groups = (Nc + cg - 1) / cg;
  const int g = c / cg;
  const int g0 = g * cg, gn = min(cg, Nc - g0);
  publish();
  const unsigned gd = wave_atomic_add(&ctl.grp_done[k * groups + g], 1u);
  if (gd != (unsigned)gn - 1u) return false;
  observe();
  if (lane == 0) ctl.grp_done[k * groups + g] = 0u;
  // the group's force configurations [g0, fe), then its others [fe, g0 + gn)
  const int fe = min(max(Nc_force, g0), g0 + gn);
  const double* b = ctl.err + (long)k * 3 * E3 + lane;
  double s[3], f[1];
  const double* const ef[1] = {b + 2 * E3 + (long)g0 * WAVE};
  sum_rows<1, VALS>(ef, WAVE, fe - g0, f);
  const double* const ee[2] = {b + (long)fe * WAVE, b + E3 + (long)fe * WAVE};
  sum_rows<2, VALS>(ee, WAVE, g0 + gn - fe, s);
  s[2] = f[0];
  if (groups == 1) {
    write_fitness(ctl, k, lane, s);
    return true;
  }
  double* gp = ctl.gpart + ((long)k * groups + g) * 3 * WAVE + lane;
  cstore(gp, s[0]);
  cstore(gp + WAVE, s[1]);
  cstore(gp + 2 * WAVE, s[2]);
  publish();
  const unsigned bd = wave_atomic_add(&ctl.blk_done[k], 1u);
  if (bd != (unsigned)groups - 1u) return false;
  observe();
  if (lane == 0) ctl.blk_done[k] = 0u;
  const double* gb = ctl.gpart + (long)k * groups * 3 * WAVE + lane;
  const double* const eg[3] = {gb, gb + WAVE, gb + 2 * WAVE};
  sum_rows<3, VALS>(eg, 3 * WAVE, groups, s);
  write_fitness(ctl, k, lane, s);
  return true;
}

// Store a chunk partial and, for the last chunk to arrive, fold all chunk
// partials of (c, k) in chunk order into acc.  Returns false for the others.
// A force configuration's partial is its force-error sum alone (acc.v[7]),
// another configuration's its energy and virial (acc.v[0..6]).
template <int VALS>
__device__ inline bool fold_chunks(const PopCtl& ctl, unsigned* counter, int c, int k, int chunk,
                                   int nch, int lane, bool force, Acc& acc)
{
  const long slot0 = ldu(ctl.cfg_slot0, c);
  double* slot = ctl.part + ((slot0 + chunk) * ctl.K + k) * 8 * WAVE + lane;
  if (force) {
    cstore(slot + 7 * WAVE, acc.v[7]);
  } else {
#pragma unroll
    for (int u = 0; u < 7; ++u) cstore(slot + u * WAVE, acc.v[u]);
  }
  publish();
  const unsigned done = wave_atomic_add(counter, 1u);
  if (done != (unsigned)nch - 1u) return false;
  observe();
  const double* base = ctl.part + (slot0 * ctl.K + k) * 8 * WAVE + lane;
  const long stride = (long)ctl.K * 8 * WAVE; // one chunk to the next
#pragma unroll
  for (int u = 0; u < 8; ++u) acc.v[u] = 0.0;
  if (force) {
    // VALS chunks' force sums per round trip, added in chunk order
    constexpr int B = VALS;
    for (int ch = 0; ch < nch; ch += B) {
      double a[B];
#pragma unroll
      for (int j = 0; j < B; ++j)
        a[j] = ch + j < nch ? cload(base + (ch + j) * stride + 7 * WAVE) : 0.0;
#pragma unroll
      for (int j = 0; j < B; ++j) acc.v[7] += a[j];
    }
  } else {
    // VALS / 7 chunks' partials per round trip (at least 3), added in chunk
    // order
    constexpr int B = VALS / 7 > 3 ? VALS / 7 : 3;
    for (int ch = 0; ch < nch; ch += B) {
      double a[B][7];
#pragma unroll
      for (int j = 0; j < B; ++j)
#pragma unroll
        for (int u = 0; u < 7; ++u)
          a[j][u] = ch + j < nch ? cload(base + (ch + j) * stride + u * WAVE) : 0.0;
#pragma unroll
      for (int j = 0; j < B; ++j)
#pragma unroll
        for (int u = 0; u < 7; ++u) acc.v[u] += a[j][u];
    }
  }
  return true;
}

// The XCD this wave runs on (hardware register; affects only which queue a
// wave serves first, never a result).
__device__ __forceinline__ unsigned xcc_id()
{
  unsigned x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
  return x & 0xfu;
}

// Next item for this wave: from queue q while it has items, then from the
// next queue (cyclically) that is still open; -1 once every queue is
// closed.  The wave that takes the last item of a queue marks it closed (one
// atomic OR per queue and launch, issued without waiting for it); a fetch
// that comes back empty reads the closed mask, so a wave finds an open
// queue, or learns that all are closed, with one load instead of trying the
// queues one by one.  (The first fetch past the end marks the queue as well,
// for a fetch that overtakes the OR.)  Counters are kFetchStride words
// apart, each in its own memory channel's lines.
__device__ inline int next_item(const PopCtl& ctl, int& q)
{
  const unsigned long long all = ctl.nq >= 64 ? ~0ull : (1ull << ctl.nq) - 1ull;
  for (;;) {
    unsigned* f = ctl.fetch + (long)q * kFetchStride;
    const unsigned q0 = (unsigned)ldu(ctl.qstart, q), len = (unsigned)ldu(ctl.qstart, q + 1) - q0;
    const unsigned it = wave_atomic_add(f, 1u);
    if (it < len) {
      if (it == len - 1u && __lane_id() == 0) atomicOr(ctl.closed, 1ull << q);
      return (int)(q0 + it);
    }
    unsigned long long closed = 0;
    if (__lane_id() == 0) {
      if (it == len) closed = atomicOr(ctl.closed, 1ull << q) | (1ull << q);
      else closed = cload(ctl.closed) | (1ull << q);
    }
    closed = ((unsigned long long)(unsigned)wave_first((int)(unsigned)(closed >> 32)) << 32) |
             (unsigned)wave_first((int)(unsigned)closed);
    const unsigned long long open = ~closed & all;
    if (open == 0) return -1;
    // the first open queue after q, else the first open one
    const unsigned long long above = q >= 63 ? 0ull : open & (~0ull << (q + 1));
    q = __builtin_ctzll(above ? above : open);
  }
}

// When several waves of a SIMD are ready, the oldest issues first, so in a
// persistent launch the youngest wave of each SIMD runs all of its items
// slowest (measured: about 1.5 us per atom for the first workgroups
// dispatched to a CU against 2.2 us for the last, with four per CU), and
// its large items end the launch.  Kernels call this once per atom with a
// running count: the wave priority then cycles through its four levels, at
// a phase set by the workgroup's dispatch round, so the waves of a SIMD take
// turns at the top.
__device__ __forceinline__ void rotate_priority(unsigned step)
{
  switch (step & 3u) {
  case 0: __builtin_amdgcn_s_setprio(0); break;
  case 1: __builtin_amdgcn_s_setprio(1); break;
  case 2: __builtin_amdgcn_s_setprio(2); break;
  default: __builtin_amdgcn_s_setprio(3); break;
  }
}

// The dispatch round of this workgroup among those sharing its CU (grids are
// sized to whole rounds of the resident workgroups).
__device__ __forceinline__ unsigned dispatch_round()
{
  return (blockIdx.x * 4u) / gridDim.x;
}

// Sum over the G lane groups of a packed item: lanes l, l ^ W, l ^ 2W, ...
// (fixed order; every lane ends with its column's total).
__device__ inline double group_sum(double x, int W)
{
  for (int m = W; m < WAVE; m <<= 1) x += __shfl_xor(x, m);
  return x;
}

// Pot must provide:
//   struct Lane; typedef Extra;                     per-lane constants, tables
//   static Lane make(ctl, p);                       denormalised genes of p
//   static void chunk(set, q, item, ctl, ex, Acc&)  evaluate atoms [a0, a1)
//   static constexpr bool kHeaderAhead;             chunk<true, LONE> exists:
//                                                   a form that loads ahead
//                                                   (used by the instantiations
//                                                   without packed items)
//   static constexpr bool kForcePass;               three-body force assembly
//   static double force_pass(set, ctl, ex, a0, a1, k, lane)
//                                                   sum |F-F_ref|^2 over [a0, a1)
//   static constexpr int kFoldValues;               values loaded per round trip
//                                                   by the reductions (registers)
//   static constexpr bool kPacked;                  packed items supported:
//   static void chunk_packed(set, q, item, ctl, ex, col, g, G, Acc&)
//   static double force_pass_packed(set, ctl, a0, a1, k, col, g, G)
//                                                   the same for lane group g
// PACK selects the instantiation that also serves packed items (chosen per
// item); the other one never contains the packed code, so its register
// budget is unaffected.  (Forcing the PACK one to the other's occupancy
// makes it spill to scratch, and a launch that needs scratch memory costs
// more than the wave it saves.)  LONE is the instantiation for launches of
// at most one wave per SIMD, where a wave's time is the latency of its
// dependent loads and registers are plentiful: with one lane block the
// individuals' constants load while the fetch is in flight, and a chunk
// loads each next atom's record during the current one and reads its rows
// from the wave's LDS copy (Pot::chunk<true, true>; only with kHeaderAhead
// and without PACK).
template <class Pot, bool PACK = false, bool LONE = false>
__global__ void __launch_bounds__(256) population_kernel(PopCtl ctl, DevSet set,
                                                         typename Pot::Extra ex)
{
  const int lane = __lane_id();
  const unsigned total_waves = gridDim.x * (blockDim.x / WAVE);
  const unsigned wave_slot = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  const int Nc = set.Nc;
  unsigned prev = 0xffffffffu;
  const unsigned long long t_entry = ctl.trace ? wall_clock64() : 0ull;
  // home queue: one of the queues of this wave's XCD (workgroups of one XCD
  // spread over its queues)
  const unsigned qpx = (unsigned)max(ctl.nq / kXcds, 1); // queues per XCD
  int q = (int)((xcc_id() * qpx + (blockIdx.x / kXcds) % qpx) % (unsigned)ctl.nq);
  // a launch with a wave for every item: each queue has as many home waves
  // as items, so every first fetch finds one
  // (wave_slot is the same in every lane, but derived from threadIdx.x the
  // compiler cannot tell: taken through readfirstlane, the load is scalar
  // and q stays in an SGPR, as do the item and everything that follows)
  const int slot_u = wave_first((int)wave_slot);
  if (slot_u < ctl.n_home) q = ldu(ctl.home, slot_u);
  bool skip_fetch = false;
  const bool early_lanes = LONE && ctl.K == 1;
  // values loaded per round trip by the reductions (more, in the lone-wave
  // form, made its launches slower: 48 values, 26.6 against 24.9 us)
  constexpr int FV = Pot::kFoldValues;
  for (;;) {
    // With one lane block, every item's individuals are known before the
    // fetch: their constants load while it is in flight.
    typename Pot::Lane lanes{};
    if (early_lanes) lanes = Pot::make(ctl, min(lane, ctl.P - 1));
    // (readfirstlane: the flag is the same in every lane, which the compiler
    // cannot tell, and a per-lane item index would move all the item's
    // values out of scalar registers)
    const int it = wave_first(skip_fetch ? -1 : next_item(ctl, q));
    unsigned long long* tr = ctl.trace ? ctl.trace + (long)kTraceFields * it : nullptr;
    if (ctl.trace && lane == 0) {
      const unsigned long long now = wall_clock64();
      if (prev != 0xffffffffu) ctl.trace[(long)kTraceFields * prev + 2] = now;
      if (it >= 0) {
        tr[0] = wave_slot;
        tr[1] = now;
        // the wave's entry time on its first item (fetch latency), else 0
        tr[3] = prev == 0xffffffffu ? t_entry : 0ull;
      }
      prev = (unsigned)it;
    }
    // marks of the item's compute phase (lane 0, recording only)
    auto mark = [&](int field) {
      if (tr && lane == 0) tr[field] = wall_clock64();
    };
    if (it < 0) {
      // Every queue is empty.  The last wave to finish returns the queue
      // counters to zero (all other waves have stopped fetching by then).
      // With one lane block, the wave that wrote the fitness is not counted:
      // it fetches nothing more, and the kernel ends with its store.
      if (!skip_fetch) {
        const unsigned d = wave_atomic_add(ctl.done, 1u);
        const unsigned counted = total_waves - (ctl.K == 1 ? 1u : 0u);
        if (d == counted - 1u && lane == 0) {
          for (int j = 0; j < ctl.nq; ++j) cstore(ctl.fetch + (long)j * kFetchStride, 0u);
          cstore(ctl.closed, 0ull);
          cstore(ctl.done, 0u);
        }
      }
      break;
    }
    const WorkItem item = ldu(ctl.items, it);
    const int c = item.c, k = item.k;
    const int nch = ldu(ctl.cfg_nchunks, c);
    const bool is_force = c < set.Nc_force;
    unsigned* cnt = &ctl.cnt[(long)c * ctl.K + k];
    Acc acc;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc.v[u] = 0.0;

    constexpr bool can_pack = PACK && Pot::kPacked;
    const bool packed = can_pack && item.groups > 1;
    if (item.phase == 0) {
      if (packed) {
        if constexpr (can_pack) {
          const int G = item.groups, W = WAVE / G;
          const int col = lane & (W - 1);
          const typename Pot::Lane q = Pot::make(ctl, min(k * WAVE + col, ctl.P - 1));
          mark(4);
          Pot::chunk_packed(set, q, item, ctl, ex, col, lane / W, G, acc);
          mark(5);
          if (!(Pot::kForcePass && is_force)) {
#pragma unroll
            for (int u = 0; u < 8; ++u) acc.v[u] = group_sum(acc.v[u], W);
          }
        }
      } else {
        // individual of this lane; lanes past P repeat the last one, write nothing
        if (!early_lanes) lanes = Pot::make(ctl, min(k * WAVE + lane, ctl.P - 1));
        mark(4);
        if constexpr (Pot::kHeaderAhead && !PACK)
          Pot::template chunk<true, LONE>(set, lanes, item, ctl, ex, acc);
        else
          Pot::chunk(set, lanes, item, ctl, ex, acc);
        mark(5);
      }
      if (Pot::kForcePass && is_force) {
        publish(); // the per-bond forces of this chunk are complete
        if (nch > 1) {
          wave_atomic_add(cnt, 1u);
          continue;
        }
        // The item was the whole configuration: this wave assembles its
        // forces at once, from its own completed stores (no assembly item,
        // no wait).
        if (packed) {
          if constexpr (can_pack) {
            const int G = item.groups, W = WAVE / G;
            acc.v[7] = group_sum(Pot::force_pass_packed(set, ctl, item.a0, item.a1, k,
                                                        lane & (W - 1), lane / W, G),
                                 W);
          }
        } else {
          acc.v[7] = Pot::force_pass(set, ctl, ex, item.a0, item.a1, k, lane);
        }
      } else if (nch > 1) {
        if (!fold_chunks<FV>(ctl, cnt, c, k, item.chunk, nch, lane, is_force, acc))
          continue;
        if (lane == 0) *cnt = 0u;
      }
    } else {
      if (!wave_wait_count(cnt, (unsigned)nch, ctl.spin_limit)) {
        if (lane == 0) atomicOr(ctl.status, 1u);
      }
      mark(4);
      if (packed) {
        if constexpr (can_pack) {
          const int G = item.groups, W = WAVE / G;
          acc.v[7] = group_sum(
            Pot::force_pass_packed(set, ctl, item.a0, item.a1, k, lane & (W - 1), lane / W, G),
            W);
        }
      } else {
        acc.v[7] = Pot::force_pass(set, ctl, ex, item.a0, item.a1, k, lane);
      }
      mark(5);
      unsigned* cnt2 = &ctl.cnt2[(long)c * ctl.K + k];
      if (nch > 1) {
        if (!fold_chunks<FV>(ctl, cnt2, c, k, item.chunk, nch, lane, true, acc))
          continue;
      }
      if (lane == 0) {
        *cnt = 0u;
        *cnt2 = 0u;
      }
    }

    double eE = 0.0, eW = 0.0, eF = 0.0;
    if (is_force) {
      eF = acc.v[7];
    } else {
      const double dE = acc.v[0] - (double)ldu(set.pe_ref, c);
      eE = dE * dE;
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        const double dW = acc.v[1 + u] - (double)ldu(set.virial_ref, c + (long)Nc * u);
        eW += dW * dW;
      }
    }
    // With one lane block, its fitness is written only once every item of
    // the launch has finished, so every queue is empty: this wave goes
    // straight to the exit (two memory round trips less at the end of the
    // launch's longest chain).
    if (config_done<FV>(ctl, Nc, set.Nc_force, c, k, lane, eE, eW, eF) &&
        ctl.K == 1)
      skip_fetch = true;
  }
}

} // namespace gpuga
