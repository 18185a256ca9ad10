/*
 * Genetic algorithm bookkeeping (test infrastructure only, see oracle.h).
 *
 * Restates src/ga.cu:
 *   constructor: cumulative rank probabilities, initial genes   :27-60
 *   insertion sort and row permutation                          :105-136
 *   crossover                                                   :161-186
 *   mutation (index drawn before value, the order g++ -std=c++14
 *   gives the unsequenced `population[r1(rng)] = r2(rng)`)       :188-198
 *   parent draw                                                 :200-212
 * The random numbers follow the C++ standard library the reference is built
 * against (libstdc++):
 *   std::mt19937           Matsumoto & Nishimura 1998, as specified by the
 *                          C++ standard [rand.eng.mers]
 *   uniform_real_distribution<float>(0,1) = generate_canonical<float,24>:
 *                          one 32-bit draw x, (float)x / 2^32, and a result
 *                          >= 1 replaced by the float just below 1
 *   uniform_int_distribution<int>(a,b): Lemire, "Fast Random Integer
 *                          Generation in an Interval", ACM TOMACS 29(1),
 *                          2019 (arXiv:1805.10941), on a 64-bit product
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_mt_seed(orc_mt19937* g, uint32_t seed)
{
  g->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->idx = 624;
}

uint32_t orc_mt_next(orc_mt19937* g)
{
  if (g->idx >= 624) {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      g->mt[i] = v;
    }
    g->idx = 0;
  }
  uint32_t y = g->mt[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

float orc_canonical_float(orc_mt19937* g)
{
  float x = (float)orc_mt_next(g);
  float r = x / 4294967296.0f;
  if (r >= 1.0f) r = nextafterf(1.0f, 0.0f);
  return r;
}

int orc_uniform_int(orc_mt19937* g, int a, int b)
{
  uint32_t range = (uint32_t)b - (uint32_t)a + 1u;
  uint64_t product = (uint64_t)orc_mt_next(g) * range;
  uint32_t low = (uint32_t)product;
  if (low < range) {
    uint32_t threshold = (uint32_t)(-range) % range;
    while (low < threshold) {
      product = (uint64_t)orc_mt_next(g) * range;
      low = (uint32_t)product;
    }
  }
  return (int)(uint32_t)(product >> 32) + a;
}

int orc_ga_init(orc_ga* ga, const orc_ga_params* p, int V, uint32_t seed)
{
  memset(ga, 0, sizeof *ga);
  ga->p = *p;
  ga->V = V;
  const int P = p->population_size;
  ga->child_number = P - p->parent_number;
  if (ga->child_number < 0 || ga->child_number % 2 || p->parent_number < 2 || V < 2) return 1;
  ga->population = (float*)malloc(sizeof(float) * P * V);
  ga->population_copy = (float*)malloc(sizeof(float) * P * V);
  ga->fitness = (float*)malloc(sizeof(float) * P);
  ga->index = (int*)malloc(sizeof(int) * P);
  ga->cumulative_probabilities = (float*)malloc(sizeof(float) * p->parent_number);
  float numerator = 0.0f;
  float denominator = (float)((1.0 + p->parent_number) * p->parent_number / 2.0);
  for (int n = 0; n < p->parent_number; ++n) {
    numerator += (float)(p->parent_number - n);
    ga->cumulative_probabilities[n] = numerator / denominator;
  }
  orc_mt_seed(&ga->rng, seed);
  for (int n = 0; n < P * V; ++n) ga->population[n] = orc_canonical_float(&ga->rng);
  return 0;
}

void orc_ga_free(orc_ga* ga)
{
  free(ga->population);
  free(ga->population_copy);
  free(ga->fitness);
  free(ga->index);
  free(ga->cumulative_probabilities);
  memset(ga, 0, sizeof *ga);
}

void orc_ga_sort(orc_ga* ga)
{
  const int P = ga->p.population_size, V = ga->V;
  float* a = ga->fitness;
  int* idx = ga->index;
  for (int n = 0; n < P; ++n) idx[n] = n;
  for (int i = 1; i < P; ++i) {
    float key = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > key) {
      a[j + 1] = a[j];
      idx[j + 1] = idx[j];
      --j;
    }
    a[j + 1] = key;
    idx[j + 1] = i;
  }
  memcpy(ga->population_copy, ga->population, sizeof(float) * P * V);
  for (int n = 0; n < P; ++n)
    memcpy(ga->population + (size_t)n * V, ga->population_copy + (size_t)idx[n] * V,
           sizeof(float) * V);
}

static int get_a_parent(orc_ga* ga)
{
  float u = orc_canonical_float(&ga->rng);
  for (int n = 0; n < ga->p.parent_number; ++n)
    if (ga->cumulative_probabilities[n] > u) return n;
  return 0;
}

void orc_ga_crossover(orc_ga* ga)
{
  const int V = ga->V;
  float* pop = ga->population;
  for (int m = 0; m < ga->child_number; m += 2) {
    int p1 = get_a_parent(ga);
    int p2 = get_a_parent(ga);
    while (p2 == p1) p2 = get_a_parent(ga);
    int cp = orc_uniform_int(&ga->rng, 1, V - 1);
    int c1 = ga->p.parent_number + m, c2 = c1 + 1;
    for (int n = 0; n < cp; ++n) {
      pop[c1 * V + n] = pop[p1 * V + n];
      pop[c2 * V + n] = pop[p2 * V + n];
    }
    for (int n = cp; n < V; ++n) {
      pop[c1 * V + n] = pop[p2 * V + n];
      pop[c2 * V + n] = pop[p1 * V + n];
    }
  }
}

void orc_ga_mutation(orc_ga* ga, int generation)
{
  const int V = ga->V, m = ga->p.population_size * V;
  float ratio = (float)generation / (float)ga->p.maximum_generation;
  float rate = (float)((double)ga->p.mutation_rate * (1.0 - (double)ratio));
  int count = (int)roundf((float)m * rate);
  for (int n = 0; n < count; ++n) {
    int idx = orc_uniform_int(&ga->rng, V, m - 1);
    float val = orc_canonical_float(&ga->rng);
    ga->population[idx] = val;
  }
}

void orc_ga_step(orc_ga* ga, int generation)
{
  orc_ga_sort(ga);
  orc_ga_crossover(ga);
  orc_ga_mutation(ga, generation);
}
