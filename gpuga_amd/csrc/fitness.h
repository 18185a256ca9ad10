// Fitness evaluator: the drop-in for the reference's class Fitness
// (src/fitness.cuh:30-83) and its Potential plug-ins (src/potential.cuh:22-41).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "ga.h"
#include "ga_device.h"
#include "kernels.h"
#include "trainset.h"

namespace gpuga {

struct FitnessOptions {
  int backend = 0;        // 0: population kernel, 1: per-individual launches
  int chunk_cost = 0;     // work-item cost target; 0 = autotune on first use
  int use_graph = 1;      // replay each generation from a captured HIP graph
  int nn2b_bf16 = 1;      // NN2B hidden layer 2 on BF16 MFMA (else FP32)
  int nn2b_grid = 1024;   // NN2B tabulation points
  int verbose = 0;
  int blocks_per_cu = 0;  // 0 = from the occupancy calculator
  int pack_lanes = 1;     // pack a last lane block of <= 16 individuals
  int queues = 0;         // work queues of the population launch; 0 = one per XCD
  int ga_batch = 0;       // generations per staged batch of the device GA; 0 = up to 32
  int tail_split = 0;     // configurations smaller than one work item are cut into
                          // pieces of 1/tail_split of an item (0, 1: not cut)
  int serve_share = 0;    // serving: 1 = each job's launches sized to its share of
                          // the GPU (by the jobs running); 0 = the whole GPU each
};

// Shape of the population launch for one population size: work-item cost,
// queue count, resident workgroups per CU and lane packing.  0 in queues or
// blocks_per_cu means the built-in choice (four queues per XCD; the
// occupancy calculator).
struct LaunchShape {
  int chunk_cost = 0;
  int queues = 0;
  int blocks_per_cu = 0;
  int pack = 1; // pack a last lane block of <= 16 individuals
};

// Device-resident copy of one training set and its geometry tables.
struct DeviceSet {
  DeviceBuffer<int> Na, Na_sum, type, nn, bond0, nl, rev, ptype, tri0;
  DeviceBuffer<float4> bond, tri, ri_tab, rec, pair_tab;
  // the Tersoff population kernel's pruned list and its tables, kept when
  // the looped kernel reads them (more than 16 neighbours within R2)
  DeviceBuffer<int> p_nn, p_b0, p_nl, p_rev, p_ptype, p_tri0;
  DeviceBuffer<float4> p_bond, p_tri;
  DeviceBuffer<int4> pair_hdr;
  DeviceBuffer<float> pe_ref, virial_ref, force_ref;
  DeviceBuffer<double> ri_cconst;
  DeviceBuffer<float> ri_fconst;
  DevSet view{};
  size_t bytes() const;
};

// Abstract plug-in with the reference's three calls (src/potential.cuh:26-40);
// pointers are device pointers laid out as the reference lays them out
// (f [3][N], virial [6][N], pe [N]).
class Potential {
public:
  virtual ~Potential() = default;
  virtual void initialize(int N, int max_Na) = 0;
  virtual void update_potential(const std::vector<float>& physical) = 0;
  // ws: device workspace of workspace_floats() floats, or null for the
  // plug-in's own
  virtual void find_force(const DevSet& set, float* f, float* virial, float* pe,
                          hipStream_t st, float* ws = nullptr) = 0;
  virtual size_t workspace_floats() const { return 0; }
};

std::unique_ptr<Potential> make_potential(const PotentialSpec& spec, const DeviceSet& ds);

// One population plan (work items, scratch) for a population size P.
struct PopPlan {
  int P = 0, K = 0, n_items = 0, grid = 0, block = 256, chunk_cost = 0;
  int occ = 0;         // resident workgroups per CU the grid is sized for
  int groups_last = 1; // lane groups of the last lane block's items
  int nq = 1;          // work queues
  DeviceBuffer<WorkItem> items;
  DeviceBuffer<int> qstart;
  DeviceBuffer<int> home; // first queue per wave slot (launches with a wave per item)
  DeviceBuffer<int> cfg_nchunks, cfg_slot0;
  DeviceBuffer<double> part, err, gpart;
  DeviceBuffer<double> out; // terms [P][3] | fitness [P] (float) | status (unsigned)
  DeviceBuffer<float> bondf, genes, nn2b_table;
  DeviceBuffer<unsigned> counters; // cnt | cnt2 | grp_done | blk_done | fetch
  PinnedBuffer<float> h_genes;
  PinnedBuffer<double> h_out;
  PopCtl ctl{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the launch on 1/2, 1/4, 1/8 of the grid (a served job sharing the GPU),
  // captured when first used
  static constexpr int kShares = 3;
  hipGraph_t share_graph[kShares] = {};
  hipGraphExec_t share_exec[kShares] = {};
  void drop_share_graphs();
  // device GA of this population size; runs of generations (population
  // launch + GA step each) as graphs: ga_exec[k][j] is 2^j generations
  // starting on population buffer k (a generation on buffer b writes the
  // next into buffer 1 - b)
  static constexpr int kGaGraphs = 6; // runs of 1, 2, 4, ..., 32 generations
  std::unique_ptr<DeviceGA> dga;
  hipGraph_t ga_graph[2][kGaGraphs] = {};
  hipGraphExec_t ga_exec[2][kGaGraphs] = {};
  void drop_ga_graphs();
  ~PopPlan();
};

class Fitness {
public:
  Fitness(const std::string& input_dir, const FitnessOptions& opt = FitnessOptions());
  Fitness(const PotentialSpec& spec, TrainSet ts, const FitnessOptions& opt = FitnessOptions());
  ~Fitness();

  // Fitness::compute (src/fitness.cu:316-333): genes [P][V] in [0,1] -> fitness[P].
  void compute(int P, const float* genes, float* fitness, double* terms = nullptr);
  // The per-individual-launch form (one find_force + reduction + sync each).
  void compute_plain(int P, const float* genes, float* fitness, double* terms = nullptr);
  // The same with every device scratch array in the fixed per-GPU workspace
  // of noalloc.h: no device memory is allocated or freed.  Fails when the
  // training set needs more than that workspace holds.
  void compute_plain_noalloc(int P, const float* genes, float* fitness, double* terms = nullptr);
  // Workspace bytes compute_plain_noalloc needs for this training set.
  size_t noalloc_bytes() const;
  // Fitness::predict (src/fitness.cu:347-390): writes force/energy/virial.out.
  void predict(const std::string& dir, const float* elite);
  // Per-atom outputs of one physical parameter set (host arrays).
  void find_force(const float* physical, float* f, float* virial, float* pe);
  // Launch the population generation asynchronously / wait for it.
  // share: the launch runs on 1/share of the plan's grid (rounded up to 1, 2,
  // 4 or 8), for a GPU shared with other jobs; results do not depend on it
  void compute_async(int P, const float* genes, int share = 1);
  void compute_wait(float* fitness, double* terms);
  // Tune the launch shape for population size P (the parts the options
  // leave open) and return its work-item cost.
  int autotune(int P);
  // The launch shape the population size P runs with (tuning it first when
  // the options ask for that).
  LaunchShape launch_shape(int P);
  // What the plan for P was built with: work-item cost, queues, resident
  // workgroups per CU, lane groups of the last lane block.
  void plan_shape(int P, int& chunk_cost, int& queues, int& blocks_per_cu, int& groups_last);
  // Build the work plan (and its graph) for population size P now, so the
  // first generation does not pay for it.
  void prepare(int P);
  // `generations` GA generations (fitness, sort, crossover, mutation) with
  // the population and the GA on the device, from generation
  // first_generation; ga holds the state before and after.  best[g] and
  // elite[g][V] (genes in [0, 1]) receive each generation's ga.out record.
  void run_ga(GA& ga, int first_generation, int generations, float* best, float* elite);
  // One generation with the item timeline recorded: rows of 12 values per
  // work item (config, lane block, chunk, atom begin, atom end, phase, wave
  // slot, start ns, end ns, entry ns of the wave for its first item else
  // NaN, compute begin and end ns; times relative to the first start).
  std::vector<double> trace(int P, const float* genes);
  // Work-item size from the training set alone (no timing; deterministic).
  int heuristic_chunk_cost(int P) const;
  // Device bytes a handle for this training set and population would hold.
  static size_t estimate_bytes(const PotentialSpec& s, const TrainSet& ts, int P,
                               const FitnessOptions& opt);

  int number_of_variables() const { return spec_.V; }
  const PotentialSpec& spec() const { return spec_; }
  const TrainSet& trainset() const { return ts_; }
  const NeighborList& neighbors() const { return nl_; }
  // The list the Tersoff population kernel's atom records hold: the bonds
  // shorter than the largest R2 any individual can have (the whole list
  // when no bond is that long).
  const NeighborList& population_neighbors() const { return pruned_ ? pnl_ : nl_; }
  // The device geometry tables (bond, triplet, atom records), copied back.
  void device_geometry(std::vector<float4>& bond, std::vector<float4>& tri,
                       std::vector<float4>& rec) const;
  hipStream_t stream() const { return stream_; }
  size_t device_bytes() const;
  size_t plan_bytes(int P) const;
  // Accumulated population-kernel time (HIP events on the launch stream).
  double kernel_ms() const { return kernel_ms_; }
  long kernel_launches() const { return kernel_launches_; }
  void reset_timers() { kernel_ms_ = 0; kernel_launches_ = 0; }
  void set_timing(bool on) { timing_ = on; }
  FitnessOptions& options() { return opt_; }
  // Polls of the population kernel's dependency wait before it reports a
  // timeout (default about a second; < 0: report one at once, for tests of
  // the failure path).
  void set_spin_limit(int spins);

private:
  void setup();
  PopPlan& plan(int P);
  void build_plan(PopPlan& pl, int P, const LaunchShape& sh);
  // the shape the options fix (chunk cost > 0: no tuning)
  LaunchShape option_shape() const;
  // the population launch on genes (default: the plan's buffer)
  void launch_population(PopPlan& pl, hipStream_t st, const float* genes = nullptr,
                         int grid = 0);
  // n generations of the device GA from population buffer k, captured
  void capture_generations(PopPlan& pl, int k, int n, hipGraph_t* graph, hipGraphExec_t* exec);
  void launch_generation(PopPlan& pl, hipStream_t st, int share = 1); // graph or direct
  void capture(PopPlan& pl, hipStream_t st);           // the plan's graph, once
  double time_plan(int P, const LaunchShape& sh, int reps, int* occ = nullptr);
  // after a launch whose status word is set: clear the plan's counters and
  // results, then fail
  [[noreturn]] void discard_launch(PopPlan& pl);
  FitnessNorm norm() const;
  NN2BGrid nn2b_grid_desc() const;
  // device arrays of the per-individual path
  struct PlainBuffers {
    float *ws, *f, *virial, *pe, *fit1;
    double *cerr, *terms1;
  };
  void plain_loop(int P, const float* genes, float* fitness, double* terms,
                  const PlainBuffers& b);
  // the per-individual path's arrays laid out from base (256-byte aligned);
  // returns the bytes they span (base may be null to measure)
  size_t carve_plain(char* base, PlainBuffers* b) const;
  // fails unless genes [P][V] suit the population kernel's tables
  void check_genes(int P, const float* genes) const;

  FitnessOptions opt_;
  PotentialSpec spec_;
  TrainSet ts_;
  NeighborList nl_;
  NeighborList pnl_;    // pruned list of the Tersoff population kernel
  bool pruned_ = false; // pnl_ holds fewer bonds than nl_
  bool loop_path_ = false; // Tersoff: more than 16 neighbours within R2
  LoopTab loop_{};         // the looped kernel's tables
  Geometry geo_; // host bond geometry of the pair potentials' set-up
  DeviceSet ds_;
  std::unique_ptr<Potential> potential_;
  std::map<int, std::unique_ptr<PopPlan>> plans_;
  std::map<int, LaunchShape> tuned_;
  hipStream_t stream_ = nullptr;
  int device_ = 0;
  int num_cu_ = 0;
  int pending_P_ = 0;
  float dmin_ = 0.0f, dmax_ = 0.0f; // pair-distance range of the training set
  // per-atom outputs and plain-path scratch
  DeviceBuffer<float> d_f_, d_virial_, d_pe_, d_params_, d_fit1_, d_pmin_, d_pmax_;
  DeviceBuffer<double> d_cerr_, d_terms1_;
  void* ws_static_ = nullptr; // this device's fixed workspace (noalloc.h)
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  bool timing_ = false;
  int spin_limit_ = kSpinLimit;
  double kernel_ms_ = 0.0;
  long kernel_launches_ = 0;
};

} // namespace gpuga
