// Genetic algorithm driver (host).  Compiled with -ffp-contract=off so the
// float arithmetic of the reference's host code is reproduced exactly.
#include "ga.h"

#include "common.h"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace gpuga {

// ------------------------------------------------------------------ RNG

void MT19937::seed_with(uint32_t seed)
{
  mt_[0] = seed;
  for (int i = 1; i < 624; ++i) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + (uint32_t)i;
  idx_ = 624;
}

void MT19937::set_state(const uint32_t* s, int idx)
{
  std::memcpy(mt_, s, sizeof mt_);
  idx_ = idx;
}

uint32_t MT19937::operator()()
{
  if (idx_ >= 624) {
    static const uint32_t mag[2] = {0u, 0x9908b0dfu};
    int i = 0;
    for (; i < 624 - 397; ++i) {
      const uint32_t y = (mt_[i] & 0x80000000u) | (mt_[i + 1] & 0x7fffffffu);
      mt_[i] = mt_[i + 397] ^ (y >> 1) ^ mag[y & 1u];
    }
    for (; i < 623; ++i) {
      const uint32_t y = (mt_[i] & 0x80000000u) | (mt_[i + 1] & 0x7fffffffu);
      mt_[i] = mt_[i + 397 - 624] ^ (y >> 1) ^ mag[y & 1u];
    }
    const uint32_t y = (mt_[623] & 0x80000000u) | (mt_[0] & 0x7fffffffu);
    mt_[623] = mt_[396] ^ (y >> 1) ^ mag[y & 1u];
    idx_ = 0;
  }
  uint32_t y = mt_[idx_++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

float MT19937::canonical()
{
  const float x = (float)(*this)();
  float r = x / 4294967296.0f;
  if (r >= 1.0f) r = std::nextafter(1.0f, 0.0f);
  return r;
}

int MT19937::uniform_int(int a, int b)
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

// --------------------------------------------------------------- ga.in

namespace {

bool valid_int(const char* s, int* out)
{
  if (!s || !*s) return false;
  char* p;
  errno = 0;
  long v = std::strtol(s, &p, 0);
  if (errno || p == s || *p) return false;
  *out = (int)v;
  return true;
}

bool valid_real(const char* s, float* out)
{
  if (!s || !*s) return false;
  char* p;
  errno = 0;
  double v = std::strtod(s, &p);
  if (errno || p == s || *p) return false;
  *out = (float)v;
  return true;
}

} // namespace

GAParams read_ga(const std::string& dir)
{
  GAParams p;
  const std::string path = dir + "/ga.in";
  std::ifstream f(path);
  if (!f) fail(strfmt("Failed to open %s!", path.c_str()));
  std::string line;
  while (std::getline(f, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ss(line);
    std::vector<std::string> tok;
    std::string t;
    while (ss >> t) tok.push_back(t);
    if (tok.empty()) continue;
    const std::string& key = tok[0];
    auto one = [&]() {
      if (tok.size() != 2) fail(key + " should have 1 parameter.\n");
      return tok[1].c_str();
    };
    if (key == "maximum_generation") {
      if (!valid_int(one(), &p.maximum_generation)) fail("maximum_generation should be integer.\n");
      if (p.maximum_generation < 1) fail("maximum_generation should be positive.\n");
    } else if (key == "population_size") {
      if (!valid_int(one(), &p.population_size)) fail("population_size should be integer.\n");
      if (p.population_size < 20) fail("population_size should >= 20.\n");
    } else if (key == "parent_number") {
      if (!valid_int(one(), &p.parent_number)) fail("parent_number should be integer.\n");
      if (p.parent_number < 10) fail("parent_number should >= 10.\n");
    } else if (key == "mutation_rate") {
      if (!valid_real(one(), &p.mutation_rate)) fail("mutation_rate should be a number.\n");
      if (p.mutation_rate < 0) fail("mutation_rate should >= 0.\n");
      if (p.mutation_rate > 1) fail("mutation_rate should <= 1.\n");
    } else if (key == "seed") {
      int s = 0;
      if (!valid_int(one(), &s)) fail("seed should be integer.\n");
      p.seed = (uint32_t)s;
    } else {
      fail("'" + key + "' is invalid keyword.\n");
    }
  }
  validate_ga(p);
  return p;
}

void validate_ga(const GAParams& p)
{
  if (p.population_size < 20) fail("population_size should >= 20.\n");
  if (p.parent_number < 10) fail("parent_number should >= 10.\n");
  if (p.parent_number >= p.population_size) fail("parent_number should < population_size.\n");
  if ((p.population_size - p.parent_number) % 2 != 0)
    fail("population_size - parent_number should be even.\n");
}

// ------------------------------------------------------------------- GA

GA::GA(const GAParams& p, int V, const std::vector<float>& pmin, const std::vector<float>& pmax)
  : p_(p), V_(V), pmin_(pmin), pmax_(pmax), rng_(p.seed)
{
  validate_ga(p);
  if (V < 2) fail("the potential needs at least 2 parameters for crossover");
  const int P = p.population_size;
  child_number_ = P - p.parent_number;
  population_.resize((size_t)P * V);
  population_copy_.resize((size_t)P * V);
  fitness_.assign(P, 0.0f);
  index_.resize(P);
  cumulative_.resize(p.parent_number);
  // src/ga.cu:41-46: float running sum over a float denominator
  float numerator = 0.0f;
  const float denominator = (float)((1.0 + p.parent_number) * p.parent_number / 2.0);
  for (int n = 0; n < p.parent_number; ++n) {
    numerator += (float)(p.parent_number - n);
    cumulative_[n] = numerator / denominator;
  }
  for (size_t n = 0; n < population_.size(); ++n) population_[n] = rng_.canonical();
}

void GA::sort_population()
{
  // src/ga.cu:105-136: stable insertion sort on the fitness array itself
  const int P = p_.population_size;
  for (int n = 0; n < P; ++n) index_[n] = n;
  float* a = fitness_.data();
  for (int i = 1; i < P; ++i) {
    const float key = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > key) {
      a[j + 1] = a[j];
      index_[j + 1] = index_[j];
      --j;
    }
    a[j + 1] = key;
    index_[j + 1] = i;
  }
  population_copy_ = population_;
  for (int n = 0; n < P; ++n)
    std::memcpy(&population_[(size_t)n * V_], &population_copy_[(size_t)index_[n] * V_],
                sizeof(float) * V_);
}

int GA::get_a_parent()
{
  // src/ga.cu:200-212
  const float u = rng_.canonical();
  for (int n = 0; n < p_.parent_number; ++n)
    if (cumulative_[n] > u) return n;
  return 0;
}

void GA::crossover()
{
  // src/ga.cu:161-186
  float* pop = population_.data();
  for (int m = 0; m < child_number_; m += 2) {
    const int p1 = get_a_parent();
    int p2 = get_a_parent();
    while (p2 == p1) p2 = get_a_parent();
    const int cp = rng_.uniform_int(1, V_ - 1);
    const int c1 = p_.parent_number + m, c2 = c1 + 1;
    for (int n = 0; n < cp; ++n) {
      pop[c1 * V_ + n] = pop[p1 * V_ + n];
      pop[c2 * V_ + n] = pop[p2 * V_ + n];
    }
    for (int n = cp; n < V_; ++n) {
      pop[c1 * V_ + n] = pop[p2 * V_ + n];
      pop[c2 * V_ + n] = pop[p1 * V_ + n];
    }
  }
}

void GA::mutation(int generation)
{
  // src/ga.cu:188-198; the gene index is drawn before the value
  const int m = p_.population_size * V_;
  const float ratio = (float)generation / (float)p_.maximum_generation;
  const float rate = (float)((double)p_.mutation_rate * (1.0 - (double)ratio));
  const int count = (int)std::round((float)m * rate);
  for (int n = 0; n < count; ++n) {
    const int idx = rng_.uniform_int(V_, m - 1);
    const float val = rng_.canonical();
    population_[idx] = val;
  }
}

void GA::write_line(FILE* fid, int generation) const
{
  std::fprintf(fid, "%d %g ", generation, fitness_[0]);
  for (int m = 0; m < V_; ++m) {
    const float a = pmin_[m];
    const float b = pmax_[m] - a;
    const float bx = b * population_[m];
    std::fprintf(fid, "%g ", a + bx);
  }
  std::fprintf(fid, "\n");
}

void GA::run(const std::string& dir, const std::function<void(int, const float*, float*)>& eval,
             int print_every)
{
  FILE* fid = nullptr;
  if (!dir.empty()) {
    const std::string path = dir + "/ga.out";
    fid = std::fopen(path.c_str(), "w");
    if (!fid) fail(strfmt("Failed to open %s!", path.c_str()));
  }
  const int P = p_.population_size;
  for (int g = 0; g < p_.maximum_generation; ++g) {
    eval(P, population_.data(), fitness_.data());
    sort_population();
    if (fid) {
      write_line(fid, g);
      std::fflush(fid);
    }
    if (print_every > 0 && (g + 1) % print_every == 0) write_line(stdout, g + 1);
    crossover();
    mutation(g);
  }
  if (fid) std::fclose(fid);
}

} // namespace gpuga

// ------------------------------------------------------------ device GA

#include "ga_device.h"

namespace gpuga {

DeviceGA::DeviceGA(const GA& ga, float* genes, int history_cap, hipStream_t st)
{
  const GAParams& p = ga.params();
  const int P = p.population_size, V = ga.V();
  if (P > kDeviceGaMaxP)
    fail(strfmt("the device GA supports populations of at most %d", kDeviceGaMaxP));
  const size_t PV = (size_t)P * V;
  if (!genes) {
    genes_own_.resize(PV);
    genes = genes_own_.get();
  }
  sorted_.resize(P);
  copy_.resize(PV);
  cum_.upload(ga.cumulative(), st);
  hist_.resize((size_t)std::max(history_cap, 1) * (1 + V));
  mt_.resize(624);
  idx_.resize(1);
  gen_.resize(1);
  last_.resize(PV);
  GPUGA_HIP(hipMemsetAsync(last_.get(), 0xff, last_.bytes(), st)); // -1
  dec_.resize(3 * (size_t)std::max((P - p.parent_number + 1) / 2, 1));
  mut_idx_.resize(PV);
  mut_val_.resize(PV);
  mut_count_.resize(1);
  d_.P = P;
  d_.V = V;
  d_.parents = p.parent_number;
  d_.children = P - p.parent_number;
  d_.max_gen = p.maximum_generation;
  d_.mutation_rate = p.mutation_rate;
  d_.genes = genes;
  d_.fitness = nullptr;
  d_.sorted = sorted_.get();
  d_.genes_copy = copy_.get();
  d_.cumulative = cum_.get();
  d_.mt = mt_.get();
  d_.mt_idx = idx_.get();
  d_.generation = gen_.get();
  d_.history = hist_.get();
  d_.history_cap = std::max(history_cap, 0);
  hist_gen0_.resize(1);
  d_.history_gen0 = hist_gen0_.get();
  d_.last = last_.get();
  d_.dec = dec_.get();
  d_.mut_idx = mut_idx_.get();
  d_.mut_val = mut_val_.get();
  d_.mut_count = mut_count_.get();
}

bool DeviceGA::matches(const GA& ga) const
{
  const GAParams& p = ga.params();
  return d_.P == p.population_size && d_.V == ga.V() && d_.parents == p.parent_number &&
         d_.max_gen == p.maximum_generation && d_.mutation_rate == p.mutation_rate;
}

void DeviceGA::upload(const GA& ga, int generation, hipStream_t st)
{
  const size_t PV = (size_t)d_.P * d_.V;
  GPUGA_HIP(hipMemcpyAsync(d_.genes, ga.population().data(), sizeof(float) * PV,
                           hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipMemcpyAsync(d_.mt, ga.rng().state(), sizeof(uint32_t) * 624,
                           hipMemcpyHostToDevice, st));
  const int idx = ga.rng().index();
  GPUGA_HIP(hipMemcpyAsync(d_.mt_idx, &idx, sizeof(int), hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipMemcpyAsync(d_.generation, &generation, sizeof(int), hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipMemcpyAsync(d_.history_gen0, &generation, sizeof(int), hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipStreamSynchronize(st)); // the host values are stack/host memory
}

void DeviceGA::restart_history(int generation, hipStream_t st)
{
  GPUGA_HIP(hipMemcpyAsync(d_.history_gen0, &generation, sizeof(int), hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipStreamSynchronize(st));
}

void DeviceGA::download(GA& ga, hipStream_t st) const
{
  const size_t PV = (size_t)d_.P * d_.V;
  uint32_t s[624];
  int idx = 0;
  GPUGA_HIP(hipMemcpyAsync(ga.population().data(), d_.genes, sizeof(float) * PV,
                           hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipMemcpyAsync(ga.fitness().data(), d_.sorted, sizeof(float) * d_.P,
                           hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipMemcpyAsync(s, d_.mt, sizeof s, hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipMemcpyAsync(&idx, d_.mt_idx, sizeof idx, hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipStreamSynchronize(st));
  ga.rng().set_state(s, idx);
}

void DeviceGA::step(const float* fitness, hipStream_t st)
{
  GaDev d = d_;
  d.fitness = fitness;
  launch_ga_generation(d, st);
}

void DeviceGA::history(float* out, int n, hipStream_t st) const
{
  if (n > d_.history_cap) fail("history beyond its capacity");
  GPUGA_HIP(hipMemcpyAsync(out, d_.history, sizeof(float) * (size_t)n * (1 + d_.V),
                           hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipStreamSynchronize(st));
}

} // namespace gpuga
