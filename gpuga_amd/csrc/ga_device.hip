// One GA generation on the device (the reference's GA::compute body after
// the fitness call: sort_population, crossover, mutation; src/ga.cu:83-88,
// 105-212), drawing the same random stream as the host GA (ga.cpp):
// std::mt19937 with libstdc++'s uniform_real_distribution<float> and
// uniform_int_distribution<int>.  Used so that a run of generations needs no
// host round trip.  Three launches:
//
//  ga_sort    (1024 threads)  stable sort of the fitness, rows reordered, the
//                             generation's ga.out record;
//  ga_decide  (one wave)      every random draw, in the host's order: the
//                             crossover decisions and the list of mutations;
//  ga_apply   (1024 threads)  crossover copies, then the mutations in list
//                             order.
//
// ga_decide keeps the 624-word state in LDS and a window of 64 consecutive
// tempered outputs in a register (one per lane); a sequential draw reads
// its lane with readlane, so a draw costs a few instructions rather than an
// LDS round trip.
#include "ga_device.h"

#include <hip/hip_runtime.h>

namespace gpuga {

namespace {

constexpr int MT_N = 624, MT_M = 397;
constexpr int GA_THREADS = 1024;
constexpr int WAVE = 64; // CDNA wavefront width

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

// float in [0, 1) from one draw as generate_canonical<float, 24> (x / 2^32
// is exact as a multiplication by 2^-32; 1.0 becomes the float below it)
__device__ __forceinline__ float canonical_of(uint32_t x)
{
  const float r = (float)x * 2.3283064365386963e-10f;
  return r >= 1.0f ? 0.99999994f : r;
}

// ---------------------------------------------------------------- sort

__global__ void __launch_bounds__(GA_THREADS) ga_sort(GaDev d)
{
  extern __shared__ unsigned char lds_raw[];
  float* fit = reinterpret_cast<float*>(lds_raw); // [P]
  int* order = reinterpret_cast<int*>(fit + d.P); // [P]
  __shared__ int s_nan;
  const int t = threadIdx.x, T = blockDim.x;
  const int P = d.P, V = d.V;
  if (t == 0) s_nan = 0;
  __syncthreads();
  for (int i = t; i < P; i += T) {
    fit[i] = d.fitness[i];
    if (fit[i] != fit[i]) s_nan = 1;
  }
  __syncthreads();
  // sort_population (src/ga.cu:105-136) is a stable insertion sort: for
  // values without NaN its result is the stable rank order; with a NaN the
  // insertion sort itself runs on one thread
  const bool nan = s_nan != 0;
  if (!nan) {
    for (int i = t; i < P; i += T) {
      const float f = fit[i];
      int r = 0;
      for (int j = 0; j < P; ++j) {
        const float g = fit[j];
        r += (g < f) || (g == f && j < i);
      }
      order[r] = i;
    }
  } else if (t == 0) {
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
  for (int n = t; n < P; n += T) d.sorted[n] = nan ? fit[n] : fit[order[n]];
  const long PV = (long)P * V;
  for (long e = t; e < PV; e += T) d.genes_copy[e] = d.genes[e];
  __syncthreads();
  for (long e = t; e < PV; e += T) {
    const long n = e / V, m = e - n * V;
    d.genes[e] = d.genes_copy[(long)order[n] * V + m];
  }
  // the ga.out record of this generation: best fitness and the elite row
  const int row = *d.generation - *d.history_gen0;
  if (d.history && row >= 0 && row < d.history_cap) {
    float* h = d.history + (long)row * (1 + V);
    if (t == 0) h[0] = nan ? fit[0] : fit[order[0]];
    for (int m = t; m < V; m += T) h[1 + m] = d.genes_copy[(long)order[0] * V + m];
  }
}

// -------------------------------------------------------------- decide

// The random stream of one wave: state words in LDS (s), tempered outputs of
// the current block in LDS (tb), and a register window of 64 of them.
struct WaveStream {
  uint32_t* s;
  uint32_t* tb;
  int idx;   // position in the block (0..624), uniform
  int base;  // first position held by the window, uniform
  uint32_t w; // tb[base + lane]

  __device__ void retemper()
  {
    for (int i = __lane_id(); i < MT_N; i += WAVE) tb[i] = temper(s[i]);
    __syncthreads();
  }
  // the state update (Matsumoto & Nishimura 1998) by one wave: word i needs
  // words i, i + 1 before the update and word i + 397 (mod 624) after it
  // when i + 397 >= 624, so three dependent ranges, each read then written
  __device__ void twist()
  {
    const int l = __lane_id();
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = l + WAVE * k;
      if (i < MT_N - MT_M) v[k] = twist_one(s[i], s[i + 1], s[i + MT_M]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = l + WAVE * k;
      if (i < MT_N - MT_M) s[i] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = MT_N - MT_M + l + WAVE * k;
      if (i < 2 * (MT_N - MT_M)) v[k] = twist_one(s[i], s[i + 1], s[i - (MT_N - MT_M)]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = MT_N - MT_M + l + WAVE * k;
      if (i < 2 * (MT_N - MT_M)) s[i] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = 2 * (MT_N - MT_M) + l + WAVE * k;
      if (i < MT_N) v[k] = twist_one(s[i], s[i + 1 < MT_N ? i + 1 : 0], s[i - (MT_N - MT_M)]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = 2 * (MT_N - MT_M) + l + WAVE * k;
      if (i < MT_N) s[i] = v[k];
    }
    __syncthreads();
    retemper();
  }
  __device__ void load_window()
  {
    base = idx;
    const int i = base + __lane_id();
    w = i < MT_N ? tb[i] : 0u;
  }
  __device__ uint32_t next()
  {
    if (idx >= MT_N) {
      twist();
      idx = 0;
      load_window();
    } else if (idx - base >= WAVE) {
      load_window();
    }
    const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)w, idx - base);
    ++idx;
    return x;
  }
  __device__ float canonical() { return canonical_of(next()); }
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
// (src/ga.cu:200-212), 64 entries per step
__device__ int parent_of(const float* cum, int parents, float u)
{
  for (int c = 0; c < parents; c += WAVE) {
    const int n = c + __lane_id();
    const unsigned long long hit = __ballot(n < parents && cum[n] > u);
    if (hit) return c + __ffsll((long long)hit) - 1;
  }
  return 0;
}

__global__ void __launch_bounds__(WAVE) ga_decide(GaDev d)
{
  extern __shared__ unsigned char lds_raw[];
  uint32_t* s = reinterpret_cast<uint32_t*>(lds_raw); // [624] state
  uint32_t* tb = s + MT_N;                              // [624] tempered
  float* cum = reinterpret_cast<float*>(tb + MT_N);     // [parents]
  __shared__ int s_first;
  const int l = __lane_id();
  const int P = d.P, V = d.V;
  for (int i = l; i < MT_N; i += WAVE) s[i] = d.mt[i];
  for (int i = l; i < d.parents; i += WAVE) cum[i] = d.cumulative[i];
  __syncthreads();
  WaveStream rs{s, tb, *d.mt_idx, 0, 0u};
  rs.retemper();
  if (rs.idx < MT_N) rs.load_window();

  // crossover (src/ga.cu:161-186): the decisions in draw order
  for (int m = 0; m < d.children; m += 2) {
    const int p1 = parent_of(cum, d.parents, rs.canonical());
    int p2 = parent_of(cum, d.parents, rs.canonical());
    while (p2 == p1) p2 = parent_of(cum, d.parents, rs.canonical());
    const int cp = rs.uniform_int(1, V - 1);
    if (l == 0) {
      d.dec[3 * (m / 2) + 0] = p1;
      d.dec[3 * (m / 2) + 1] = p2;
      d.dec[3 * (m / 2) + 2] = cp;
    }
  }

  // mutation (src/ga.cu:188-198): count with the host's float arithmetic,
  // then gene index and value per mutation.  The mutations that fit in the
  // rest of the state block use draws idx + 2j and idx + 2j + 1; those before
  // the first rejected gene-index draw are listed at once, that one is
  // redrawn in order.
  const int gen = *d.generation;
  const int mtot = P * V;
  const float ratio = (float)gen / (float)d.max_gen;
  const float rate = (float)((double)d.mutation_rate * (1.0 - (double)ratio));
  const int count = (int)roundf((float)mtot * rate);
  const uint32_t range = (uint32_t)(mtot - 1) - (uint32_t)V + 1u;
  const uint32_t threshold = (0u - range) % range;
  int n = 0;
  while (n < count) {
    if (rs.idx >= MT_N) {
      rs.twist();
      rs.idx = 0;
      rs.load_window();
    }
    const int avail = (MT_N - rs.idx) / 2;
    const int batch = min(avail, count - n);
    if (l == 0) s_first = batch;
    __syncthreads();
    for (int j = l; j < batch; j += WAVE) {
      const uint64_t m = (uint64_t)tb[rs.idx + 2 * j] * range;
      if ((uint32_t)m < threshold) atomicMin(&s_first, j);
    }
    __syncthreads();
    const int ok = s_first;
    for (int j = l; j < ok; j += WAVE) {
      const uint64_t m = (uint64_t)tb[rs.idx + 2 * j] * range;
      d.mut_idx[n + j] = (int)(uint32_t)(m >> 32) + V;
      d.mut_val[n + j] = canonical_of(tb[rs.idx + 2 * j + 1]);
    }
    const int left = MT_N - rs.idx - 2 * ok; // draws left in the block
    rs.idx += 2 * ok;
    n += ok;
    if (n < count && (ok < batch || left == 1)) {
      // a rejected gene-index draw, or a mutation across the block end
      if (rs.idx < MT_N) rs.load_window();
      const int idx = rs.uniform_int(V, mtot - 1);
      const float val = rs.canonical();
      if (l == 0) {
        d.mut_idx[n] = idx;
        d.mut_val[n] = val;
      }
      ++n;
    }
    if (rs.idx < MT_N) rs.load_window();
  }

  for (int i = l; i < MT_N; i += WAVE) d.mt[i] = s[i];
  if (l == 0) {
    *d.mt_idx = rs.idx;
    *d.mut_count = count;
    *d.generation = gen + 1;
  }
}

// --------------------------------------------------------------- apply

__global__ void __launch_bounds__(GA_THREADS) ga_apply(GaDev d)
{
  const int t = threadIdx.x, T = blockDim.x;
  const int V = d.V;
  // parents are rows < parents, children rows >= parents: no overlap
  const long CV = (long)d.children * V;
  for (long e = t; e < CV; e += T) {
    const int c = (int)(e / V), n = (int)(e - (long)c * V);
    const int* q = d.dec + 3 * (c / 2);
    const int a = (c & 1) ? q[1] : q[0], b = (c & 1) ? q[0] : q[1];
    d.genes[(long)(d.parents + c) * V + n] = d.genes[(long)(n < q[2] ? a : b) * V + n];
  }
  __syncthreads();
  // mutations in list order: the last one on a gene wins (d.last holds -1
  // outside this step)
  const int count = *d.mut_count;
  for (int n = t; n < count; n += T) atomicMax(&d.last[d.mut_idx[n]], n);
  __syncthreads();
  for (int n = t; n < count; n += T)
    if (__hip_atomic_load(&d.last[d.mut_idx[n]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == n)
      d.genes[d.mut_idx[n]] = d.mut_val[n];
  __syncthreads();
  for (int n = t; n < count; n += T) d.last[d.mut_idx[n]] = -1;
}

} // namespace

size_t ga_device_lds_bytes(int P, int parents, int)
{
  const size_t sort = sizeof(float) * P + sizeof(int) * P;
  const size_t decide = sizeof(uint32_t) * 2 * MT_N + sizeof(float) * parents;
  return sort > decide ? sort : decide;
}

void launch_ga_generation(const GaDev& d, hipStream_t st)
{
  ga_sort<<<1, GA_THREADS, sizeof(float) * d.P + sizeof(int) * d.P, st>>>(d);
  ga_decide<<<1, WAVE, sizeof(uint32_t) * 2 * MT_N + sizeof(float) * d.parents, st>>>(d);
  ga_apply<<<1, GA_THREADS, 0, st>>>(d);
  GPUGA_HIP(hipGetLastError());
}

} // namespace gpuga
