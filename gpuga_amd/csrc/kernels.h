// Host-callable launchers of the HIP kernels (definitions in the .hip files).
#pragma once

#include "common.h"
#include "device.h"
#include "population.h"

namespace gpuga {

// ---- single parameter set (Potential::find_force semantics, per-atom out)
void launch_tersoff_plain(const DevSet& s, const TersoffParams& q, float* b, float* bp,
                          float* f12, float* f, float* virial, float* pe, hipStream_t st);
// tab: per-bond RI factors (r^-6, erfc(a r)/r, -(d/dr) erfc(a r)/r, 1/r)
void launch_ri_plain(const DevSet& s, const RIParams& q, const float4* tab, float* f,
                     float* virial, float* pe, hipStream_t st);
// w: device pointer to the n(n+4)+1 physical parameters (species: then the
// n weights of the species-pair input)
void launch_nn2b_plain(const DevSet& s, int n, bool species, const float* w, float* f,
                       float* virial, float* pe, hipStream_t st);

// ---- fitness of one parameter set from per-atom outputs
// cerr: [Nc][3] scratch; out: fitness (float) and terms[3] (double)
void launch_fitness_terms(const DevSet& s, const FitnessNorm& nrm, const float* f,
                          const float* virial, const float* pe, double* cerr, float* fitness,
                          double* terms, hipStream_t st);

// ---- population kernels (one persistent launch per generation)
// cls: neighbour-count class of the training set (0: <= 4, 1: <= 8, 2: <= 16);
// pack: serve the packed-item queue (several atoms per wave); lone: a launch
// of at most one wave per SIMD (the form that trades registers for latency)
void launch_tersoff_population(const PopCtl& ctl, const DevSet& s, int cls, int grid, int block,
                               bool pack, bool lone, hipStream_t st);
// Tersoff sets with more than 16 neighbours within R2: per-bond lane values
// in LDS (dynamic, tersoff_loop_lds_bytes per workgroup), geometry from the
// bond and triplet tables of the list in t.
struct LoopTab {
  const int *nn, *b0, *tri0, *rev, *ptype; // the list the kernel evaluates
  const float4 *bond, *tri;                // its bond and triplet tables
  int max_nn;
};
size_t tersoff_loop_lds_bytes(int max_nn, int block);
void launch_tersoff_loop_population(const PopCtl& ctl, const DevSet& s, const LoopTab& t,
                                    int grid, int block, hipStream_t st);
int tersoff_loop_occupancy(bool two, int max_nn, int block);
void launch_ri_population(const PopCtl& ctl, const DevSet& s, const float4* ri_tab, int grid,
                          int block, hipStream_t st);

// ---- NN2B population: tabulate U on a grid per individual (MFMA), then the
// pair kernel interpolates
struct NN2BGrid {
  int G;          // grid points
  float d0, h;    // first point and spacing
  int n;          // neurons
  int bf16;       // 1: hidden layer 2 on BF16 MFMA, 0: FP32 VALU
  int codes;      // species codes tabulated (1; 3 with the species-pair input)
};
void launch_nn2b_tabulate(const PopCtl& ctl, const NN2BGrid& g, float* table, hipStream_t st);
void launch_nn2b_population(const PopCtl& ctl, const DevSet& s, const NN2BGrid& g,
                            const float* table, int grid, int block, hipStream_t st);

// Per-kernel register/occupancy information for the launch planner.
int population_max_blocks_per_cu(int potential_type, int block);

} // namespace gpuga
