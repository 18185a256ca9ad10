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

void MT19937::twist()
{
  // the state update of std::mt19937 (C++ [rand.eng.mers]), branch-free
  int i = 0;
  for (; i < 624 - 397; ++i) {
    const uint32_t y = (mt_[i] & 0x80000000u) | (mt_[i + 1] & 0x7fffffffu);
    mt_[i] = mt_[i + 397] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  }
  for (; i < 623; ++i) {
    const uint32_t y = (mt_[i] & 0x80000000u) | (mt_[i + 1] & 0x7fffffffu);
    mt_[i] = mt_[i + 397 - 624] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  }
  const uint32_t y = (mt_[623] & 0x80000000u) | (mt_[0] & 0x7fffffffu);
  mt_[623] = mt_[396] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  idx_ = 0;
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

void GA::draw_crossover(std::vector<int>& cross)
{
  // src/ga.cu:161-186: two parents (distinct), then the crossover point
  for (int m = 0; m < child_number_; m += 2) {
    const int p1 = get_a_parent();
    int p2 = get_a_parent();
    while (p2 == p1) p2 = get_a_parent();
    const int cp = rng_.uniform_int(1, V_ - 1);
    cross.push_back(p1);
    cross.push_back(p2);
    cross.push_back(cp);
  }
}

void GA::crossover()
{
  cross_.clear();
  draw_crossover(cross_);
  float* pop = population_.data();
  for (int m = 0; m < child_number_; m += 2) {
    const int* q = &cross_[(size_t)3 * (m / 2)];
    const int p1 = q[0], p2 = q[1], cp = q[2];
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

int GA::mutation_count(int generation) const
{
  // src/ga.cu:190-192 with the same float/double arithmetic
  const int m = p_.population_size * V_;
  const float ratio = (float)generation / (float)p_.maximum_generation;
  const float rate = (float)((double)p_.mutation_rate * (1.0 - (double)ratio));
  return std::max(0, (int)std::round((float)m * rate));
}

int GA::draw_mutation(int generation, std::vector<int>& idx, std::vector<float>& val)
{
  // src/ga.cu:193-197; the gene index is drawn before the value
  const int m = p_.population_size * V_;
  const int count = mutation_count(generation);
  for (int n = 0; n < count; ++n) {
    idx.push_back(rng_.uniform_int(V_, m - 1));
    val.push_back(rng_.canonical());
  }
  return count;
}

void GA::mutation(int generation)
{
  mut_idx_.clear();
  mut_val_.clear();
  const int count = draw_mutation(generation, mut_idx_, mut_val_);
  for (int n = 0; n < count; ++n) population_[mut_idx_[n]] = mut_val_[n];
}

void write_ga_line(FILE* fid, int generation, float best, const float* genes,
                   const std::vector<float>& pmin, const std::vector<float>& pmax)
{
  std::fprintf(fid, "%d %g ", generation, best);
  for (size_t m = 0; m < pmin.size(); ++m) {
    const float a = pmin[m];
    const float b = pmax[m] - a;
    const float bx = b * genes[m];
    std::fprintf(fid, "%g ", a + bx);
  }
  std::fprintf(fid, "\n");
}

void GA::write_line(FILE* fid, int generation) const
{
  write_ga_line(fid, generation, fitness_[0], population_.data(), pmin_, pmax_);
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

#include <algorithm>
#include <climits>

namespace gpuga {

DeviceGA::DeviceGA(const GA& ga, float* genes0, int history_cap, hipStream_t st, int max_batch)
{
  const GAParams& p = ga.params();
  const int P = p.population_size, V = ga.V();
  if (P > kDeviceGaMaxP)
    fail(strfmt("the device GA supports populations of at most %d", kDeviceGaMaxP));
  ga_step_prepare(P, V);
  const size_t PV = (size_t)P * V;
  if (!genes0) {
    genes_own_.resize(PV);
    genes0 = genes_own_.get();
  }
  genes0_ = genes0;
  genes1_.resize(PV);
  sorted_.resize(P);
  hist_.resize((size_t)std::max(history_cap, 1) * (1 + V));
  gen_.resize(1);
  hist_gen0_.resize(1);
  counters_.resize(2); // steps | workgroups done
  counters_.zero(st);
  // a record: the mutation count, the crossover words, the mutations' gene
  // indices and values; at most the mutations of generation 0 (the count
  // falls with the generation)
  pairs_ = (P - p.parent_number) / 2;
  max_mut_ = ga.mutation_count(0);
  int rows = 0;
  size_t lds = 0;
  ga_step_shape(P, V, &rows, &lds);
  const size_t rec_max = 1 + 3 * (size_t)pairs_ + 2 * (size_t)max_mut_;
  const size_t budget = (size_t)1 << 22; // words per batch
  const size_t most = max_batch > 0 ? std::min(max_batch, kGaMaxBatch) : kGaMaxBatch;
  const int batch = (int)std::max<size_t>(1, std::min<size_t>(most, budget / rec_max));
  stage_words_ = 1 + (size_t)batch + (size_t)batch * rec_max;
  if (stage_words_ > (size_t)INT_MAX) fail("device GA: draw records too large");
  draws_.resize(stage_words_);
  for (int k = 0; k < 2; ++k) {
    host_[k].resize(stage_words_);
    GPUGA_HIP(hipEventCreateWithFlags(&copied_[k], hipEventDisableTiming));
  }
  d_.P = P;
  d_.V = V;
  d_.parents = p.parent_number;
  d_.children = P - p.parent_number;
  d_.rows = rows;
  d_.genes_in = genes0_;
  d_.genes_out = genes1_.get();
  d_.fitness = nullptr;
  d_.sorted = sorted_.get();
  d_.draws = draws_.get();
  d_.batch = batch;
  d_.step = counters_.get();
  d_.done = counters_.get() + 1;
  d_.generation = gen_.get();
  d_.history = hist_.get();
  d_.history_cap = std::max(history_cap, 0);
  d_.history_gen0 = hist_gen0_.get();
  max_gen_ = p.maximum_generation;
  rate_ = p.mutation_rate;
  GPUGA_HIP(hipStreamSynchronize(st));
}

DeviceGA::~DeviceGA()
{
  for (int k = 0; k < 2; ++k)
    if (copied_[k]) (void)hipEventDestroy(copied_[k]);
}

bool DeviceGA::matches(const GA& ga) const
{
  const GAParams& p = ga.params();
  return d_.P == p.population_size && d_.V == ga.V() && d_.parents == p.parent_number &&
         max_gen_ == p.maximum_generation && rate_ == p.mutation_rate;
}

void DeviceGA::upload(const GA& ga, int generation, hipStream_t st)
{
  if (generation < 0) fail("generation numbers start at 0");
  const size_t PV = (size_t)d_.P * d_.V;
  GPUGA_HIP(hipMemcpyAsync(genes0_, ga.population().data(), sizeof(float) * PV,
                           hipMemcpyHostToDevice, st));
  const unsigned zero[2] = {0u, 0u};
  GPUGA_HIP(hipMemcpyAsync(d_.step, zero, sizeof zero, hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipMemcpyAsync(d_.generation, &generation, sizeof(int), hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipMemcpyAsync(d_.history_gen0, &generation, sizeof(int), hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipStreamSynchronize(st)); // the host values are stack/host memory
  staged_ = 0;
}

void DeviceGA::restart_history(int generation, hipStream_t st)
{
  GPUGA_HIP(hipMemcpyAsync(d_.history_gen0, &generation, sizeof(int), hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipStreamSynchronize(st));
}

void DeviceGA::stage(GA& ga, int generation, int n, hipStream_t st)
{
  const int B = d_.batch;
  if (n < 1 || n > B) fail(strfmt("device GA: %d records staged, batches hold 1-%d", n, B));
  if (generation < 0) fail("generation numbers start at 0");
  const int k = next_;
  next_ ^= 1;
  // the staging area's previous copy has run (it was queued two batches ago)
  if (pending_[k]) GPUGA_HIP(hipEventSynchronize(copied_[k]));
  int* h = host_[k].get();
  h[0] = (int)staged_; // the step that reads the first record
  size_t w = (size_t)B + 1;
  for (int i = 0; i < n; ++i) {
    h[1 + i] = (int)w;
    cross_.clear();
    idx_.clear();
    val_.clear();
    ga.draw_crossover(cross_);
    const int count = ga.draw_mutation(generation + i, idx_, val_);
    if (count > max_mut_) fail("device GA: more mutations than the staging area holds");
    // in draw order; the step keeps the last one of each gene, as the
    // reference's sequential writes do
    int* r = h + w;
    r[0] = count;
    std::memcpy(r + 1, cross_.data(), sizeof(int) * 3 * (size_t)pairs_);
    std::memcpy(r + 1 + 3 * pairs_, idx_.data(), sizeof(int) * (size_t)count);
    std::memcpy(r + 1 + 3 * pairs_ + count, val_.data(), sizeof(float) * (size_t)count);
    w += 1 + 3 * (size_t)pairs_ + 2 * (size_t)count;
  }
  // In pieces of 16 KiB: one host-to-device copy of a few tens of KiB
  // queued behind a running batch held the host until the device was done
  // with that batch, or longer (6 ms measured for 27 KB with four
  // generations queued, 32 KiB pieces as well; 8 and 16 KiB pieces return
  // at once), and the device then waited for the next batch's launch.  (A
  // piece costs about 5 us of host time: 0.15 ms for a generation of the
  // NN2B workload's 59 000 mutations.)
  constexpr size_t piece = 4096; // words
  for (size_t o = 0; o < w; o += piece)
    GPUGA_HIP(hipMemcpyAsync(draws_.get() + o, h + o, sizeof(int) * std::min(piece, w - o),
                             hipMemcpyHostToDevice, st));
  GPUGA_HIP(hipEventRecord(copied_[k], st));
  pending_[k] = true;
  staged_ += n;
}

void DeviceGA::step(int k, const float* fitness, hipStream_t st) const
{
  GaDev d = d_;
  d.genes_in = genes(k);
  d.genes_out = genes(1 - k);
  d.fitness = fitness;
  launch_ga_step(d, st);
}

void DeviceGA::download(GA& ga, int k, hipStream_t st) const
{
  const size_t PV = (size_t)d_.P * d_.V;
  GPUGA_HIP(hipMemcpyAsync(ga.population().data(), genes(k), sizeof(float) * PV,
                           hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipMemcpyAsync(ga.fitness().data(), d_.sorted, sizeof(float) * d_.P,
                           hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipStreamSynchronize(st));
}

void DeviceGA::history(float* out, int n, hipStream_t st) const
{
  if (n > d_.history_cap) fail("history beyond its capacity");
  GPUGA_HIP(hipMemcpyAsync(out, d_.history, sizeof(float) * (size_t)n * (1 + d_.V),
                           hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipStreamSynchronize(st));
}

} // namespace gpuga
