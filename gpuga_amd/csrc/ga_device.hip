// One GA generation on the device (the reference's GA::compute body after
// the fitness call: sort_population, crossover, mutation; src/ga.cu:83-88,
// 105-212) in one workgroup, drawing the same random stream as the host GA
// (ga.cpp): std::mt19937 with libstdc++'s uniform_real_distribution<float>
// and uniform_int_distribution<int>.  Used so that a run of generations
// needs no host round trip.
//
// Every thread runs the GA's sequential decisions redundantly (same values,
// same branches), so the cooperative steps (the Mersenne-Twister state
// update, barriers) can sit inside them.  Work without a data dependence
// (ranking the fitness values, copying rows, mutations drawn from one state
// block) is spread over the threads.
#include "ga_device.h"

#include <hip/hip_runtime.h>

namespace gpuga {

namespace {

constexpr int MT_N = 624, MT_M = 397;
constexpr int GA_THREADS = 1024;

__device__ __forceinline__ uint32_t temper(uint32_t y)
{
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

__device__ __forceinline__ uint32_t twist_one(uint32_t a, uint32_t b, uint32_t c)
{
  const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
  return c ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

// The state update of all 624 words (Matsumoto & Nishimura 1998) by the
// whole block.  Word i needs words i, i + 1 before the update and word
// i + 397 (mod 624) after it when i + 397 >= 624, so it runs in three
// dependent ranges: [0, 227), [227, 454), [454, 624); each range reads its
// inputs, waits, then writes.
__device__ void mt_twist(uint32_t* s)
{
  const int t = threadIdx.x;
  // [0, 227): all inputs old
  uint32_t v = 0;
  if (t < MT_N - MT_M) v = twist_one(s[t], s[t + 1], s[t + MT_M]);
  __syncthreads();
  if (t < MT_N - MT_M) s[t] = v;
  __syncthreads();
  // [227, 454): s[i - 227] new
  const int i2 = t + (MT_N - MT_M);
  if (i2 < 2 * (MT_N - MT_M)) v = twist_one(s[i2], s[i2 + 1], s[i2 - (MT_N - MT_M)]);
  __syncthreads();
  if (i2 < 2 * (MT_N - MT_M)) s[i2] = v;
  __syncthreads();
  // [454, 624): word 623 pairs with the new word 0
  const int i3 = t + 2 * (MT_N - MT_M);
  if (i3 < MT_N) v = twist_one(s[i3], s[i3 + 1 < MT_N ? i3 + 1 : 0], s[i3 - (MT_N - MT_M)]);
  __syncthreads();
  if (i3 < MT_N) s[i3] = v;
  __syncthreads();
}

// The stream: state words in LDS, position idx held identically by every
// thread.  All calls are made by all threads with the same idx.
struct Stream {
  uint32_t* s;
  int idx;
  __device__ uint32_t next()
  {
    if (idx >= MT_N) {
      mt_twist(s);
      idx = 0;
    }
    return temper(s[idx++]);
  }
  // generate_canonical<float, 24> on one 32-bit draw; 1.0 becomes the float
  // below it (x / 2^32 is exact as a multiplication by 2^-32)
  __device__ float canonical()
  {
    const float r = (float)next() * 2.3283064365386963e-10f;
    return r >= 1.0f ? 0.99999994f : r;
  }
  // uniform_int_distribution<int>(a, b): Lemire's method on a 64-bit product
  __device__ int uniform_int(int a, int b)
  {
    const uint32_t range = (uint32_t)b - (uint32_t)a + 1u;
    uint64_t m = (uint64_t)next() * range;
    uint32_t low = (uint32_t)m;
    if (low < range) {
      const uint32_t threshold = (0u - range) % range;
      while (low < threshold) {
        m = (uint64_t)next() * range;
        low = (uint32_t)m;
      }
    }
    return (int)(uint32_t)(m >> 32) + a;
  }
};

// first n with cum[n] > u (cum is non-decreasing), 0 when there is none
// (src/ga.cu:200-212)
__device__ int parent_of(const float* cum, int parents, float u)
{
  int lo = 0, hi = parents;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cum[mid] > u) hi = mid;
    else lo = mid + 1;
  }
  return lo < parents ? lo : 0;
}

__global__ void __launch_bounds__(GA_THREADS) ga_generation(GaDev d)
{
  extern __shared__ unsigned char lds_raw[];
  uint32_t* mt = reinterpret_cast<uint32_t*>(lds_raw);             // [624]
  float* cum = reinterpret_cast<float*>(mt + MT_N);                // [parents]
  float* fit = cum + d.parents;                                    // [P]
  int* order = reinterpret_cast<int*>(fit + d.P);                  // [P]
  int* dec = order + d.P;                                          // [3 * children / 2]
  __shared__ int s_nan, s_first;

  const int t = threadIdx.x, T = blockDim.x;
  const int P = d.P, V = d.V;
  const int gen = *d.generation;
  for (int i = t; i < MT_N; i += T) mt[i] = d.mt[i];
  for (int i = t; i < d.parents; i += T) cum[i] = d.cumulative[i];
  if (t == 0) s_nan = 0;
  __syncthreads();
  for (int i = t; i < P; i += T) {
    fit[i] = d.fitness[i];
    if (fit[i] != fit[i]) s_nan = 1;
  }
  __syncthreads();

  // ---- sort_population (src/ga.cu:105-136): stable insertion sort.  For
  // values without NaN its result is the stable rank order; with a NaN the
  // insertion sort itself runs on one thread.
  if (!s_nan) {
    for (int i = t; i < P; i += T) {
      const float f = fit[i];
      int r = 0;
      for (int j = 0; j < P; ++j) {
        const float g = fit[j];
        r += (g < f) || (g == f && j < i);
      }
      order[r] = i;
    }
    __syncthreads();
  } else {
    if (t == 0) {
      for (int n = 0; n < P; ++n) order[n] = n;
      for (int i = 1; i < P; ++i) {
        const float key = fit[i];
        int j = i - 1;
        while (j >= 0 && fit[j] > key) {
          fit[j + 1] = fit[j];
          order[j + 1] = order[j];
          --j;
        }
        fit[j + 1] = key;
        order[j + 1] = i;
      }
    }
    __syncthreads();
  }
  // sorted fitness (the reference sorts its fitness array in place)
  float sorted0 = 0.0f;
  if (!s_nan) sorted0 = fit[order[0]];
  else sorted0 = fit[0];
  for (int n = t; n < P; n += T) d.sorted[n] = s_nan ? fit[n] : fit[order[n]];
  // rows in sorted order
  const long PV = (long)P * V;
  for (long e = t; e < PV; e += T) d.genes_copy[e] = d.genes[e];
  __syncthreads();
  for (long e = t; e < PV; e += T) {
    const long n = e / V, m = e - n * V;
    d.genes[e] = d.genes_copy[(long)order[n] * V + m];
  }
  __syncthreads();
  // the ga.out record of this generation: best fitness and the elite row
  const int row = gen - d.history_gen0;
  if (d.history && row >= 0 && row < d.history_cap) {
    float* h = d.history + (long)row * (1 + V);
    if (t == 0) h[0] = sorted0;
    for (int m = t; m < V; m += T) h[1 + m] = d.genes[m];
  }

  // ---- crossover (src/ga.cu:161-186): decisions in draw order, copies after
  Stream rs{mt, *d.mt_idx};
  for (int m = 0; m < d.children; m += 2) {
    const int p1 = parent_of(cum, d.parents, rs.canonical());
    int p2 = parent_of(cum, d.parents, rs.canonical());
    while (p2 == p1) p2 = parent_of(cum, d.parents, rs.canonical());
    const int cp = rs.uniform_int(1, V - 1);
    if (t == 0) {
      dec[3 * (m / 2) + 0] = p1;
      dec[3 * (m / 2) + 1] = p2;
      dec[3 * (m / 2) + 2] = cp;
    }
  }
  __syncthreads();
  // parents are rows < parents, children rows >= parents: no overlap
  const long CV = (long)d.children * V;
  for (long e = t; e < CV; e += T) {
    const int c = (int)(e / V), n = (int)(e - (long)c * V);
    const int* q = dec + 3 * (c / 2);
    const int a = (c & 1) ? q[1] : q[0], b = (c & 1) ? q[0] : q[1];
    d.genes[(long)(d.parents + c) * V + n] = d.genes[(long)(n < q[2] ? a : b) * V + n];
  }
  __syncthreads();

  // ---- mutation (src/ga.cu:188-198), float arithmetic as the host's
  const int mtot = P * V;
  const float ratio = (float)gen / (float)d.max_gen;
  const float rate = (float)((double)d.mutation_rate * (1.0 - (double)ratio));
  const int count = (int)roundf((float)mtot * rate);
  const uint32_t range = (uint32_t)(mtot - 1) - (uint32_t)V + 1u;
  const uint32_t threshold = (0u - range) % range;
  int n = 0;
  while (n < count) {
    // mutations that fit in the current state block without a rejected draw
    // use draws idx + 2j (gene index) and idx + 2j + 1 (value)
    if (rs.idx >= MT_N) {
      mt_twist(mt);
      rs.idx = 0;
    }
    const int avail = (MT_N - rs.idx) / 2;
    const int batch = min(avail, count - n);
    if (t == 0) s_first = batch;
    __syncthreads();
    for (int j = t; j < batch; j += T) {
      const uint64_t m = (uint64_t)temper(mt[rs.idx + 2 * j]) * range;
      if ((uint32_t)m < threshold) atomicMin(&s_first, j);
    }
    __syncthreads();
    const int ok = s_first; // mutations [0, ok) of the batch need no redraw
    // later mutations of the batch overwrite earlier ones on the same gene:
    // the last writer per gene wins (d.last holds -1 outside this step)
    for (int j = t; j < ok; j += T) {
      const uint64_t m = (uint64_t)temper(mt[rs.idx + 2 * j]) * range;
      atomicMax(&d.last[(int)(uint32_t)(m >> 32) + V], j);
    }
    __syncthreads();
    for (int j = t; j < ok; j += T) {
      const uint64_t m = (uint64_t)temper(mt[rs.idx + 2 * j]) * range;
      const int idx = (int)(uint32_t)(m >> 32) + V;
      if (d.last[idx] == j) {
        const float r = (float)temper(mt[rs.idx + 2 * j + 1]) * 2.3283064365386963e-10f;
        d.genes[idx] = r >= 1.0f ? 0.99999994f : r;
      }
    }
    __syncthreads();
    for (int j = t; j < ok; j += T) {
      const uint64_t m = (uint64_t)temper(mt[rs.idx + 2 * j]) * range;
      d.last[(int)(uint32_t)(m >> 32) + V] = -1;
    }
    __syncthreads();
    rs.idx += 2 * ok;
    n += ok;
    if (n < count && ok < batch) {
      // the batch's first redraw, or a mutation across the block end: in
      // draw order, one at a time
      const int idx = rs.uniform_int(V, mtot - 1);
      const float val = rs.canonical();
      if (t == 0) d.genes[idx] = val;
      __syncthreads();
      ++n;
    } else if (n < count && ok == batch && batch == avail && avail * 2 < MT_N - rs.idx + 2 * ok) {
      // an odd draw left at the block end: this mutation spans two blocks
      const int idx = rs.uniform_int(V, mtot - 1);
      const float val = rs.canonical();
      if (t == 0) d.genes[idx] = val;
      __syncthreads();
      ++n;
    }
  }

  for (int i = t; i < MT_N; i += T) d.mt[i] = mt[i];
  if (t == 0) {
    *d.mt_idx = rs.idx;
    *d.generation = gen + 1;
  }
}

} // namespace

size_t ga_device_lds_bytes(int P, int parents, int children)
{
  return sizeof(uint32_t) * MT_N + sizeof(float) * parents + sizeof(float) * P +
         sizeof(int) * P + sizeof(int) * 3 * ((children + 1) / 2);
}

void launch_ga_generation(const GaDev& d, hipStream_t st)
{
  const size_t lds = ga_device_lds_bytes(d.P, d.parents, d.children);
  ga_generation<<<1, GA_THREADS, lds, st>>>(d);
  GPUGA_HIP(hipGetLastError());
}

} // namespace gpuga
