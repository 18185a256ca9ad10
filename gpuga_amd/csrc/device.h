// Device-side views of a resident training set and the shared kernel helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "params.h"

namespace gpuga {

constexpr int WAVE = 64; // CDNA wavefront width

// Read-only training set in HBM: per-atom SoA arrays, CSR neighbour list and
// the parameter-independent geometry tables.
// doubles per pair type of a configuration's RI sums (ri_cconst), floats
// per force atom (ri_fconst)
constexpr int kRiTypeConst = 22;
constexpr int kRiAtomConst = 27;

struct DevSet {
  int Nc, Nc_force, N, N_force;
  const int* Na;          // [Nc]
  const int* Na_sum;      // [Nc]
  const int* type;        // [N]
  const int* nn;          // [N] neighbour count
  const int* bond0;       // [N] first bond of atom i
  const int* nl;          // [B] neighbour atom of bond b
  const int* rev;         // [B] reverse bond of bond b
  const int* ptype;       // [B] type_i + type_j of bond b
  const float4* bond;     // [B] r_ij.x, r_ij.y, r_ij.z, |r_ij|
  const int* tri0;        // [N] first triplet of atom i (Tersoff only)
  const float4* tri;      // [T] cos theta, d cos / d r_ij (Tersoff only)
  // Tersoff atom records, one per atom at a fixed stride of 1 + M + M*M +
  // M/2 float4 (M = neighbour bound of the set): header {nn, unlike-type bit
  // mask, first bond, longest bond's length}, then M bonds (x, y, z, d),
  // then the M x M triplets (cos theta_jik, d cos / d r_ij), then M
  // reverse-bond indices (int) and M inverse distances 1/d.  Every load of
  // an atom is independent of the others, so they issue together.
  const float4* rec;
  int rec_nmax;
  int bond_force;         // bonds of the force configurations (they come first)
  // Pair-potential population tables.  An atom's pairs are all its bonds in
  // a force configuration and those with j > i otherwise.
  //  RI:   per atom {first float4, pairs of type 0, 1, 2} (grouped by
  //        type t_i + t_j); a force configuration's pair is one float4
  //        (unit vector r_ij / d, d), another configuration's two: (d,
  //        u_x u_x d, u_y u_y d, u_z u_z d), (u_x u_y d, u_y u_z d,
  //        u_z u_x d, 0).  The terms that are a parameter times a function
  //        of d alone come summed over the geometry: ri_cconst, ri_fconst.
  //  NN2B: per atom {first pair, pairs, 0, 0} (ordered by d); per pair three
  //        float4: (unit vector, d), cubic Hermite weights on the tabulation
  //        grid, (fc, fc', grid cell as int bits, species code t_i + t_j as
  //        int bits with the species-pair input, else 0).
  const int4* pair_hdr;
  const float4* pair_tab;
  // RI sums over the pairs an energy configuration evaluates, per pair type
  // t (kRiTypeConst doubles each): n, sum d, sum erfc(a d)/d, sum d^-6, then
  // per virial component (xx, yy, zz, xy, yz, zx) sum v, sum v 6 d^-7,
  // sum v h, with v = u_a u_b d and h = -(d/dr)(erfc(a d)/d)  [Nc][3][22];
  // and over a force atom's bonds, per type and component x: sum u_x 6 d^-7,
  // sum u_x, sum u_x h  [N_force][3][3][3]
  const double* ri_cconst;
  const float* ri_fconst;
  const float* pe_ref;    // [Nc]
  const float* virial_ref;// [6][Nc]
  const float* force_ref; // [3][N]
};

// Normalisers and weights of the fitness (src/fitness.cu:329-331).
struct FitnessNorm {
  double inv_pss, inv_vss, inv_fss;
  double w_energy, w_stress, w_force;
};

// Genes in [0,1] to physical parameters, float arithmetic with no fused
// multiply-add, exactly as src/fitness.cu:319-324.
__device__ inline float denormalise(float lo, float hi, float x)
{
#pragma clang fp contract(off)
  float b = hi - lo;
  float bx = b * x;
  return lo + bx;
}

// Broadcast lane 0's value to the whole wavefront.
__device__ inline int wave_first(int v) { return __builtin_amdgcn_readfirstlane(v); }

} // namespace gpuga
