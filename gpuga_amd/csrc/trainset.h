// Training set, potential description and parameter-independent geometry.
//
// Host-side mirror of the reference's set-up layer:
//   potential.in reader        src/fitness.cu:220-314
//   train.in reader            src/fitness.cu:59-218
//   box inverse                src/fitness.cu:40-57
//   neighbour list             src/neighbor.cu:30-83, src/mic.cuh:20-32
// plus the tables the population kernels read instead of recomputing the
// minimum-image geometry per individual (the reference recomputes it in every
// kernel, src/minimal_tersoff.cu:150-153 and :162-167).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace gpuga {

// POT_NN2B_SPECIES (no reference counterpart; the reference rejects type 4):
// NN2B whose first layer also takes the pair's species code s = t_i + t_j,
// x1 = tanh(w0 r + ws s - b0), with the n weights ws after the reference's
// n (n + 4) + 1 parameters.  Types must be 0 and 1.
enum PotentialType {
  POT_RI = 0,
  POT_TERSOFF1 = 1,
  POT_TERSOFF2 = 2,
  POT_NN2B = 3,
  POT_NN2B_SPECIES = 4
};

struct PotentialSpec {
  int type = POT_TERSOFF1;
  int num_neurons = 0;
  int V = 0;                 // number_of_variables
  float cutoff = 0.0f;       // neighbour-list cutoff
  float w_force = 0.0f, w_energy = 0.0f, w_stress = 0.0f;
  std::vector<float> pmin, pmax;
  std::vector<std::string> names;
  bool nn2b() const { return type == POT_NN2B || type == POT_NN2B_SPECIES; }
  // species codes a pair can have (NN2B with the species input: 0, 1, 2)
  int pair_codes() const { return type == POT_NN2B_SPECIES ? 3 : 1; }
};

struct TrainSet {
  int Nc = 0, Nc_force = 0, N = 0, N_force = 0, max_Na = 0;
  std::vector<int> Na, Na_sum, type;
  std::vector<float> h;          // [Nc][18]: box columns a, b, c, then inverse
  std::vector<float> r;          // [3][N]
  std::vector<float> pe_ref;     // [Nc]
  std::vector<float> virial_ref; // [6][Nc]
  std::vector<float> force_ref;  // [3][N]
  // Normalisers of the three relative RMS errors, accumulated in FP64.
  double energy_minimum = 0, force_square_sum = 0, potential_square_sum = 0,
         virial_square_sum = 0;
};

// CSR neighbour list: atom i's neighbours are NL[start[i] .. start[i]+NN[i]),
// ascending; rev[b] is the bond index of the reverse bond (the slot of i in
// j's list, or j's first slot if absent, as src/minimal_tersoff.cu:368-375).
struct NeighborList {
  std::vector<int> NN, start, NL, rev;
  int total = 0;
  int max_nn = 0;
};

struct Vec4 {
  float x, y, z, w;
};

// Parameter-independent geometry, computed once per training set in FP64 from
// the FP32 inputs and stored in FP32.
struct Geometry {
  std::vector<Vec4> bond;     // per bond: r_ij (x, y, z) and |r_ij|
  std::vector<int> tri_start; // per atom: offset of its NN x NN triplet block
  std::vector<Vec4> tri;      // per (i, s, t): cos(theta_jik), d cos / d r_ij
  long long tri_total = 0;
};

PotentialSpec read_potential(const std::string& dir);
TrainSet read_train(const std::string& dir);
// Builds a training set from arrays (box9 = file-order box rows per config).
TrainSet make_train(int Nc, int Nc_force, const int* Na, const int* type, const float* box9,
                    const float* r, const float* pe_ref, const float* virial_ref,
                    const float* force_ref);
void box_inverse(float* h18);
void finish_normalisers(TrainSet& ts);

NeighborList build_neighbors(const TrainSet& ts, float cutoff);
// The same list built on the current HIP device (neighbor.hip), bit-identical.
NeighborList build_neighbors_device(const TrainSet& ts, float cutoff, hipStream_t st);
// Its per-atom neighbour counts alone.
std::vector<int> count_neighbors_device(const TrainSet& ts, float cutoff, hipStream_t st);
Geometry build_geometry(const TrainSet& ts, const NeighborList& nl, bool triplets);
// Offsets of each atom's NN x NN triplet block and their total.
std::vector<int> triplet_offsets(const NeighborList& nl, long long* total);
// Per-bond pair type: three-body potentials 1 when the species differ (else
// 0), pair potentials t_i + t_j.
std::vector<int> bond_types(const TrainSet& ts, const NeighborList& nl, bool three_body);
// Bond slots M (4, 8 or 16) of the fixed-stride atom records of the Tersoff
// population kernel, 0 when an atom has more than 16 neighbours; a record
// holds 1 + M + M*M + M/2 float4: header (nn, species-pair bit mask, first
// bond as ints, the longest bond's length as a float), M bonds, M x M
// triplets (s, t), M reverse-bond ints and M inverse distances, zero-padded.
int record_slots(int max_nn);
// Largest outer cut-off radius R2 any individual of a minimal-Tersoff
// potential can have: the last parameter's pmin + (pmax - pmin) g at g = 1,
// in the float arithmetic of the denormalisation (monotonic in g, so every
// g in [0, 1] gives at most this); +inf for other potentials.
float tersoff_reach(const PotentialSpec& s);
// The bonds of nl with |r_ij| < reach (d[b] = |r_b| in FP32, as the
// geometry table holds it).  A bond at or beyond every individual's R2 has
// fc = fc' = 0 (src/minimal_tersoff.cu:77-90), so every term it enters is
// an exact zero: energy, forces and virial are unchanged without it.  A
// bond's reverse has the same length (exact minimum image), so the pruned
// list is symmetric; reverse indices are renumbered.
NeighborList prune_neighbors(const NeighborList& nl, const std::vector<float>& d, float reach);
std::vector<Vec4> build_records(const NeighborList& nl, const Geometry& g,
                                const std::vector<int>& ptype, int M);

// The device forms (neighbor.hip), bit-identical to the host ones above.
struct DeviceNeighbors {
  const int *NN = nullptr, *start = nullptr, *NL = nullptr, *rev = nullptr, *tri0 = nullptr;
  int total = 0;
};
// bond[total]; tri[triplet total] when tri != nullptr (needs d.tri0)
void build_geometry_device(const TrainSet& ts, const DeviceNeighbors& d, float4* bond,
                           float4* tri, hipStream_t st);
void build_records_device(int N, int M, const DeviceNeighbors& d, const int* ptype,
                          const float4* bond, const float4* tri, float4* rec, hipStream_t st);

// Index of the configuration that holds atom i.
int config_of(const TrainSet& ts, int i);

} // namespace gpuga
