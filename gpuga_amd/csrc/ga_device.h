// The GA's generation step on the device (ga_device.hip) and the host object
// that holds its state next to the population buffers.
#pragma once

#include <cstdint>
#include <vector>

#include "common.h"

namespace gpuga {

class GA;

// Device view of the GA state for one generation step (all device pointers).
//
// Draw records: the random decisions of one generation, drawn on the host
// (GA::draw_crossover / GA::draw_mutation) and staged in a batch buffer of
// at most `batch` records: word 0 is the step (counted from the upload) that
// reads the first record, words [1, batch] the record offsets (in words,
// from the buffer start), and a record is
//   [0]                      n = mutations of the generation
//   [1, 1 + 3 * pairs)       crossover: parent rank p1, p2, crossover point
//   [.., + n)                gene index of each mutation, in draw order
//   [.., + n)                its value (float bits)
// (A gene drawn more than once keeps its last value; the step resolves
// that on the device.)
// Step s reads record s - word 0 of the buffer staged before it.
struct GaDev {
  int P, V, parents, children;
  int rows;                // population rows per workgroup
  const float* genes_in;   // [P][V] population of this generation, in [0, 1]
  float* genes_out;        // [P][V] the next generation
  const float* fitness;    // [P] fitness of genes_in
  float* sorted;           // [P] fitness in sorted order (out)
  const int* draws;        // batch buffer of draw records
  int batch;               // records per batch
  unsigned* step;          // [1] steps since the upload
  unsigned* done;          // [1] workgroups finished, 0 between steps
  int* generation;         // [1] generation number, advanced by each step
  float* history;          // [history_cap][1 + V] best fitness and elite row
  int history_cap;         // rows; generation g goes to row g - *history_gen0
  int* history_gen0;       // [1]
};

// Largest population the step supports (LDS budget of its sort).
constexpr int kDeviceGaMaxP = 4096;
// Most draw records in one staged batch.
constexpr int kGaMaxBatch = 32;

// rows per workgroup and LDS bytes of the step for shape (P, V)
void ga_step_shape(int P, int V, int* rows, size_t* lds);
// checks the shape fits a workgroup and raises the LDS limit when needed
// (once per shape, outside graph capture)
void ga_step_prepare(int P, int V);
void launch_ga_step(const GaDev& d, hipStream_t st);

// Device state of a host GA for runs of generations without host round
// trips: two population buffers (a step reads one and writes the other),
// the sorted fitness, the ga.out history, the staged draw records.  Buffer 0
// may be an external device buffer (the population kernel's input).
class DeviceGA {
public:
  // max_batch: most generations per staged batch (0: up to kGaMaxBatch)
  DeviceGA(const GA& ga, float* genes0, int history_cap, hipStream_t st, int max_batch = 0);
  ~DeviceGA();
  DeviceGA(const DeviceGA&) = delete;
  DeviceGA& operator=(const DeviceGA&) = delete;
  // whether this state was built for a GA of ga's shape and settings
  bool matches(const GA& ga) const;
  // host population -> buffer 0, generation number, step count 0; the
  // history restarts at that generation
  void upload(const GA& ga, int generation, hipStream_t st);
  // the history restarts at `generation` (for runs longer than its capacity)
  void restart_history(int generation, hipStream_t st);
  int history_cap() const { return d_.history_cap; }
  int batch() const { return d_.batch; }
  // Draw the records of `n` <= batch() generations from generation
  // `generation` on the host GA (advancing its random stream as the host
  // steps would) and queue their copy to the device on st.  The n steps
  // that read them must be queued after this call and before the next
  // stage().
  void stage(GA& ga, int generation, int n, hipStream_t st);
  // one step reading population buffer k and writing buffer 1 - k, on
  // fitness (device pointer); graph-capturable
  void step(int k, const float* fitness, hipStream_t st) const;
  // population buffer k and sorted fitness -> host GA (its random stream is
  // already where the staged steps leave it)
  void download(GA& ga, int k, hipStream_t st) const;
  // history rows [n][1 + V] of the generations since the last restart
  void history(float* out, int n, hipStream_t st) const;
  float* genes(int k) const { return k ? genes1_.get() : genes0_; }

private:
  GaDev d_{};
  int pairs_ = 0, max_mut_ = 0, max_gen_ = 0;
  float rate_ = 0.0f;
  size_t stage_words_ = 0;
  long staged_ = 0; // records staged since the upload
  float* genes0_ = nullptr;
  DeviceBuffer<float> genes_own_, genes1_, sorted_, hist_;
  DeviceBuffer<int> draws_, gen_, hist_gen0_;
  DeviceBuffer<unsigned> counters_;
  // two host staging areas, each reusable once its copy has run
  PinnedBuffer<int> host_[2];
  hipEvent_t copied_[2] = {nullptr, nullptr};
  bool pending_[2] = {false, false};
  int next_ = 0;
  std::vector<int> cross_, idx_;
  std::vector<float> val_;
};

} // namespace gpuga
