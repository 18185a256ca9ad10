// Genetic algorithm driver: the drop-in for the reference's class GA
// (src/ga.cuh:20-58, src/ga.cu:27-316).  The random numbers reproduce the
// libstdc++ <random> stream the reference draws from, with explicit
// algorithms so that the sequence does not depend on the host library.
#pragma once

#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

namespace gpuga {

class Fitness;

// std::mt19937 (Matsumoto & Nishimura 1998, C++ [rand.eng.mers]).
class MT19937 {
public:
  explicit MT19937(uint32_t seed = 5489u) { seed_with(seed); }
  void seed_with(uint32_t seed);
  // (inline: the GA draws hundreds of thousands per generation)
  uint32_t operator()()
  {
    if (idx_ >= 624) twist();
    uint32_t y = mt_[idx_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  // uniform_real_distribution<float>(0, 1) of libstdc++: generate_canonical
  // with one 32-bit draw, (float)x / 2^32, results >= 1 replaced by the float
  // just below 1.
  float canonical()
  {
    const float r = (float)(*this)() * 2.3283064365386963e-10f; // exact: 2^-32
    return r >= 1.0f ? 0.99999994f : r;                          // nextafter(1, 0)
  }
  // uniform_int_distribution<int>(a, b) of libstdc++ >= 11: Lemire's
  // nearly-divisionless method on a 64-bit product (ACM TOMACS 29(1), 2019).
  int uniform_int(int a, int b)
  {
    const uint32_t range = (uint32_t)b - (uint32_t)a + 1u;
    uint64_t m = (uint64_t)(*this)() * range;
    uint32_t low = (uint32_t)m;
    if (low < range) {
      const uint32_t threshold = (uint32_t)(0u - range) % range;
      while (low < threshold) {
        m = (uint64_t)(*this)() * range;
        low = (uint32_t)m;
      }
    }
    return (int)(uint32_t)(m >> 32) + a;
  }
  const uint32_t* state() const { return mt_; }
  int index() const { return idx_; }
  void set_state(const uint32_t* s, int idx);

private:
  void twist();
  uint32_t mt_[624];
  int idx_;
};

struct GAParams {
  int maximum_generation = 1000; // src/ga.cuh:29-34 defaults
  int population_size = 200;
  int parent_number = 100;
  float mutation_rate = 0.2f;
  uint32_t seed = 12345678u;     // the reference's deterministic seed (src/ga.cu:47-48)
};

// ga.in reader (src/ga.cu:214-316 with the tokenizer of src/read_file.cu:46-72).
// Extension: an optional `seed` keyword.  Population and parent counts need
// not be multiples of 10; the child count must be even.
GAParams read_ga(const std::string& dir);
void validate_ga(const GAParams& p);

// One ga.out line (src/ga.cu:138-147): generation, best fitness, then the
// physical parameters a + (b - a) x of the elite genes x, each "%g ".
void write_ga_line(FILE* fid, int generation, float best, const float* genes,
                   const std::vector<float>& pmin, const std::vector<float>& pmax);

class GA {
public:
  GA(const GAParams& p, int V, const std::vector<float>& pmin, const std::vector<float>& pmax);
  // one generation after fitness() has been filled: sort, crossover, mutation
  void sort_population();
  void crossover();
  void mutation(int generation);
  // The random decisions of crossover and mutation, drawn from the stream in
  // the order crossover() and mutation() draw them (they use these).  They
  // depend on the stream and the generation number, not on the fitness, so a
  // run can draw them ahead of the evaluations.
  // crossover (src/ga.cu:161-186): per child pair, the parent ranks p1, p2
  // and the crossover point, appended to cross
  void draw_crossover(std::vector<int>& cross);
  // mutation (src/ga.cu:188-198): gene indices and values in draw order,
  // appended to idx and val; returns their number
  int draw_mutation(int generation, std::vector<int>& idx, std::vector<float>& val);
  // number of mutations of a generation
  int mutation_count(int generation) const;
  int children() const { return child_number_; }
  // the reference's ga.out line (src/ga.cu:141-147)
  void write_line(FILE* fid, int generation) const;
  // full run as GA::GA + GA::compute (src/ga.cu:27-92); fitness callback
  // evaluates P genes rows into fitness.  Writes <dir>/ga.out when dir != "".
  void run(const std::string& dir, const std::function<void(int, const float*, float*)>& eval,
           int print_every = 10);

  std::vector<float>& population() { return population_; }
  std::vector<float>& fitness() { return fitness_; }
  const std::vector<float>& population() const { return population_; }
  const std::vector<float>& cumulative() const { return cumulative_; }
  MT19937& rng() { return rng_; }
  const MT19937& rng() const { return rng_; }
  int V() const { return V_; }
  const GAParams& params() const { return p_; }

private:
  int get_a_parent();
  GAParams p_;
  int V_;
  int child_number_;
  std::vector<float> pmin_, pmax_;
  std::vector<float> population_, population_copy_, fitness_, cumulative_;
  std::vector<int> index_;
  std::vector<int> cross_, mut_idx_;
  std::vector<float> mut_val_;
  MT19937 rng_;
};

} // namespace gpuga
