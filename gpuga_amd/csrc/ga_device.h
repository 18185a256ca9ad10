// The GA's generation step on the device (ga_device.hip) and the host object
// that holds its state next to a population buffer.
#pragma once

#include <cstdint>
#include <vector>

#include "common.h"

namespace gpuga {

class GA;

// Device view of the GA state for one generation step (all device pointers).
struct GaDev {
  int P, V, parents, children, max_gen;
  float mutation_rate;
  float* genes;            // [P][V] population in [0, 1], updated in place
  const float* fitness;    // [P] fitness of the population
  float* sorted;           // [P] fitness in sorted order (out)
  float* genes_copy;       // [P][V] scratch
  const float* cumulative; // [parents] rank-selection probabilities
  uint32_t* mt;            // [624] std::mt19937 state words
  int* mt_idx;             // [1] position in the state block
  int* generation;         // [1] generation number, advanced by each step
  float* history;          // [history_cap][1 + V] best fitness and elite row
  int history_cap;         // rows; generation g goes to row g - *history_gen0
  int* history_gen0;       // [1]
  int* last;               // [P][V] scratch, -1 between steps
  int* dec;                // [children / 2][3] crossover parents and point
  int* mut_idx;            // [P * V] mutation list: gene index ...
  float* mut_val;          // [P * V] ... and value, in draw order
  int* mut_count;          // [1]
};

// Largest population the one-workgroup step supports (LDS budget).
constexpr int kDeviceGaMaxP = 4096;

size_t ga_device_lds_bytes(int P, int parents, int children);
void launch_ga_generation(const GaDev& d, hipStream_t st);

// Device copy of a host GA (population, random state, selection table) for a
// run of generations without host round trips.  genes may be an external
// device buffer (the population kernel's input) or owned here.
class DeviceGA {
public:
  DeviceGA(const GA& ga, float* genes, int history_cap, hipStream_t st);
  // whether this state was built for a GA of ga's shape and settings
  bool matches(const GA& ga) const;
  // host GA -> device (population, random state, generation number); the
  // history restarts at that generation
  void upload(const GA& ga, int generation, hipStream_t st);
  // the history restarts at `generation` (for runs longer than its capacity)
  void restart_history(int generation, hipStream_t st);
  int history_cap() const { return d_.history_cap; }
  // device -> host GA (population, random state, sorted fitness)
  void download(GA& ga, hipStream_t st) const;
  // one generation step on fitness (device pointer)
  void step(const float* fitness, hipStream_t st);
  // history rows [n][1 + V] of the generations since upload
  void history(float* out, int n, hipStream_t st) const;
  const GaDev& view() const { return d_; }
  float* genes() const { return d_.genes; }

private:
  GaDev d_{};
  DeviceBuffer<float> genes_own_, sorted_, copy_, cum_, hist_, fit_, mut_val_;
  DeviceBuffer<uint32_t> mt_;
  DeviceBuffer<int> idx_, gen_, last_, dec_, mut_idx_, mut_count_, hist_gen0_;
};

} // namespace gpuga
