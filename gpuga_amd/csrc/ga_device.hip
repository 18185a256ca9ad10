// One GA generation on the device: the reference's GA::compute body after
// the fitness call (sort_population, crossover, mutation; src/ga.cu:83-88,
// 105-198) as one launch.
//
// The random decisions (parent ranks, crossover points, mutated genes and
// their values) depend on the random stream and the generation number only,
// never on the fitness, so the host draws them ahead of the run with the
// host GA's own stream (GA::draw_crossover / GA::draw_mutation) and stages
// them as draw records; a record's mutations arrive with overwritten ones
// removed (the last mutation of a gene wins, as in the reference's loop).
// What depends on the fitness is done here:
//
//   sort      stable order of the fitness (the result of the reference's
//             insertion sort): by rank for P <= 512, by a bitonic sort of
//             (fitness, index) pairs above; with a NaN present the insertion
//             sort itself runs on one thread, since NaN makes the
//             reference's order depend on the insertion sequence;
//   rows      each workgroup builds `rows` rows of the next population in
//             LDS (parents: the sorted rows; children: crossover of two
//             sorted parents), applies the record's mutations that fall in
//             them, and stores the rows; workgroup 0 also writes the sorted
//             fitness and the generation's ga.out record.
//
// The step reads one population buffer and writes the other, so no
// workgroup waits for another; the last one to finish advances the step and
// generation counters.
#include "ga_device.h"

#include <hip/hip_runtime.h>

namespace gpuga {

namespace {

constexpr int GA_THREADS = 1024;
constexpr int RANK_MAX_P = 512;
constexpr int TILE_FLOATS = 4096; // rows of the next population per workgroup

template <class T>
__device__ __forceinline__ T aload(const T* p)
{
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ __forceinline__ void astore(T* p, T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int pow2_at_least(int n)
{
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

__global__ void __launch_bounds__(GA_THREADS) ga_step(GaDev d)
{
  extern __shared__ unsigned char lds_raw[];
  const int P = d.P, V = d.V;
  const int Ppad = pow2_at_least(P);
  const int pairs = d.children >> 1;
  float* fit = reinterpret_cast<float*>(lds_raw); // [Ppad] fitness as given
  float* key = fit + Ppad;                        // [Ppad] fitness, sorted
  int* ord = reinterpret_cast<int*>(key + Ppad);  // [Ppad] index of each sorted entry
  int* cross = ord + Ppad;                        // [3 * pairs] the record's crossover
  float* tile = reinterpret_cast<float*>(cross + 3 * pairs);
  __shared__ int s_nan, s_gen, s_gen0;
  __shared__ unsigned s_step;
  __shared__ int s_off[kGaMaxBatch + 1];
  const int t = threadIdx.x, T = blockDim.x;

  // counters, the batch's record offsets and the fitness in one round trip
  if (t == 0) {
    s_nan = 0;
    s_step = aload(d.step);
    s_gen = aload(d.generation);
    s_gen0 = aload(d.history_gen0);
  }
  if (t <= d.batch) s_off[t] = d.draws[t];
  for (int i = t; i < Ppad; i += T) {
    // padding sorts after every entry: +inf with an index >= P
    const float f = i < P ? d.fitness[i] : __builtin_inff();
    fit[i] = f;
    key[i] = f;
    ord[i] = i;
    if (f != f) s_nan = 1;
  }
  __syncthreads();
  const unsigned step = s_step;
  // word 0: the step of the batch's first record; then the offsets (the
  // host stages exactly the steps it queues; the clamp only keeps a
  // misuse inside the buffer)
  const int r = min(max((int)(step - (unsigned)s_off[0]), 0), d.batch - 1);
  const int* rec = d.draws + s_off[1 + r];

  // ---- sort_population (src/ga.cu:105-136), with the record's crossover
  // words copied to LDS beside it
  if (s_nan) {
    for (int i = t; i < 3 * pairs; i += T) cross[i] = rec[1 + i];
    if (t == 0) {
      for (int i = 1; i < P; ++i) {
        const float k = key[i];
        int j = i - 1;
        while (j >= 0 && key[j] > k) {
          key[j + 1] = key[j];
          ord[j + 1] = ord[j];
          --j;
        }
        key[j + 1] = k;
        ord[j + 1] = i;
      }
    }
  } else if (P <= RANK_MAX_P) {
    // stable rank: entries before i are smaller, or equal with a smaller
    // index; the upper half of the workgroup loads the crossover words
    if (t < RANK_MAX_P) {
      for (int i = t; i < P; i += RANK_MAX_P) {
        const float f = fit[i];
        int r = 0;
        for (int j = 0; j < P; ++j) {
          const float g = fit[j];
          r += (g < f) || (g == f && j < i);
        }
        key[r] = f;
        ord[r] = i;
      }
    } else {
      for (int i = t - RANK_MAX_P; i < 3 * pairs; i += T - RANK_MAX_P) cross[i] = rec[1 + i];
    }
  } else {
    for (int i = t; i < 3 * pairs; i += T) cross[i] = rec[1 + i];
    // bitonic sort of (fitness, index): distinct keys, so the result is the
    // stable order
    for (int k = 2; k <= Ppad; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = t; i < Ppad; i += T) {
          const int l = i ^ j;
          if (l > i) {
            const float a = key[i], b = key[l];
            const int ia = ord[i], ib = ord[l];
            const bool after = a > b || (a == b && ia > ib);
            if (after == ((i & k) == 0)) {
              key[i] = b;
              key[l] = a;
              ord[i] = ib;
              ord[l] = ia;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  __syncthreads();

  if (blockIdx.x == 0) {
    for (int n = t; n < P; n += T) d.sorted[n] = key[n];
    // the ga.out record of this generation (src/ga.cu:138-147): best
    // fitness and the best row
    const int row = s_gen - s_gen0;
    if (d.history && row >= 0 && row < d.history_cap) {
      float* h = d.history + (long)row * (1 + V);
      if (t == 0) h[0] = key[0];
      for (int m = t; m < V; m += T) h[1 + m] = d.genes_in[(long)ord[0] * V + m];
    }
  }

  // ---- rows of the next population: sorted parents, then crossover
  // (src/ga.cu:161-186), then mutation (src/ga.cu:188-198)
  const int r0 = blockIdx.x * d.rows;
  const int r1 = min(P, r0 + d.rows);
  const int nel = (r1 - r0) * V;
  const long e0 = (long)r0 * V;
  // the generation's mutations in draw order: gene indices, then values
  const int nm = rec[0];
  const int* mi = rec + 1 + 3 * pairs;
  const float* mv = reinterpret_cast<const float*>(mi + nm);
  // per gene of this workgroup's rows, the last mutation that hits it (the
  // reference applies them in draw order, so the last one stays)
  int* win = reinterpret_cast<int*>(tile + (size_t)d.rows * V);
  for (int e = t; e < nel; e += T) win[e] = -1;
  for (int e = t; e < nel; e += T) {
    const int n = r0 + e / V, m = e % V;
    int src;
    if (n < d.parents) {
      src = ord[n];
    } else {
      const int c = n - d.parents;
      const int* q = cross + 3 * (c >> 1);
      const int a = (c & 1) ? q[1] : q[0], b = (c & 1) ? q[0] : q[1];
      src = ord[m < q[2] ? a : b];
    }
    tile[e] = d.genes_in[(long)src * V + m];
  }
  __syncthreads();
  // every workgroup scans all the mutations, four indices per thread and
  // round trip
  for (int j = 4 * t; j < nm; j += 4 * T) {
    int g[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) g[u] = j + u < nm ? mi[j + u] : -1;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long i = (long)g[u] - e0;
      if (g[u] >= 0 && i >= 0 && i < nel) atomicMax(&win[i], j + u);
    }
  }
  __syncthreads();
  for (int e = t; e < nel; e += T)
    if (win[e] >= 0) tile[e] = mv[win[e]];
  __syncthreads();
  for (int e = t; e < nel; e += T) d.genes_out[e0 + e] = tile[e];

  // the last workgroup advances the counters every workgroup read above
  __syncthreads();
  if (t == 0) {
    const unsigned prev =
      __hip_atomic_fetch_add(d.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == gridDim.x - 1) {
      astore(d.step, step + 1u);
      astore(d.generation, s_gen + 1);
      astore(d.done, 0u);
    }
  }
}

} // namespace

// rows per workgroup and LDS bytes of a step of shape (P, V)
void ga_step_shape(int P, int V, int* rows, size_t* lds)
{
  int Ppad = 1;
  while (Ppad < P) Ppad <<= 1;
  const int r = V >= TILE_FLOATS ? 1 : TILE_FLOATS / V;
  *rows = r;
  const size_t pairs = (size_t)(P / 2); // at most: children < P
  // sort keys and order, crossover words, the rows' tile and, per gene of
  // the tile, the last mutation that hits it
  *lds = (sizeof(float) * 2 + sizeof(int)) * (size_t)Ppad + sizeof(int) * 3 * pairs +
         (sizeof(float) + sizeof(int)) * (size_t)r * V;
}

void ga_step_prepare(int P, int V)
{
  int rows = 0;
  size_t lds = 0;
  ga_step_shape(P, V, &rows, &lds);
  if (lds > 160u * 1024u)
    fail(strfmt("device GA: %d parameters need more LDS than a workgroup has", V));
  if (lds > 64u * 1024u)
    GPUGA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&ga_step),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

void launch_ga_step(const GaDev& d, hipStream_t st)
{
  int rows = 0;
  size_t lds = 0;
  ga_step_shape(d.P, d.V, &rows, &lds);
  if (rows != d.rows) fail("device GA: row split differs from its plan");
  const int grid = (d.P + rows - 1) / rows;
  ga_step<<<grid, GA_THREADS, lds, st>>>(d);
  GPUGA_HIP(hipGetLastError());
}

} // namespace gpuga
