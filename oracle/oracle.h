/*
 * FP64 host oracle for the GPUGA fitness-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against; it is never linked into, loaded by or called
 * from the product library, its CLI or its Python package.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * It restates, in scalar C, the mathematics of the reference program
 * (brucefan1983/GPUGA, file:line citations on every function):
 *   - input parsing          src/fitness.cu:59-314, src/ga.cu:214-316,
 *                            src/read_file.cu:46-102
 *   - box inverse, MIC, NL   src/fitness.cu:40-57, src/mic.cuh:20-32,
 *                            src/neighbor.cu:30-66   (FP32, bit-exact form)
 *   - minimal Tersoff 1/2    src/minimal_tersoff.cu:39-405   (FP64)
 *   - RI                     src/ri.cu:36-187                (FP64)
 *   - NN2B                   src/nn2b.cu:33-171              (FP64)
 *   - NN2B + species input   type 4, no reference counterpart (FP64; see pair.c)
 *   - fitness reduction      src/fitness.cu:316-333,437-555  (FP64)
 *   - genetic algorithm      src/ga.cu:27-212 with libstdc++ <random>
 *                            semantics (mt19937, generate_canonical<float>,
 *                            uniform_int_distribution via Lemire's method)
 *
 * Pinning status of each function is listed in the test files that check it
 * (tests/test_oracle_*.py).  Every function here is pinned by at least one of:
 * reference-held output vectors, closed forms, finite differences,
 * invariants, or the C++ standard library's <random>.
 */
#ifndef GPUGA_ORACLE_H
#define GPUGA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- inputs */

typedef struct {
  int potential_type;  /* 0 RI, 1 Tersoff-1, 2 Tersoff-2, 3 NN2B, 4 NN2B with the
                          species-pair input */
  int num_neurons;     /* NN2B only */
  int num_variables;
  float cutoff;        /* neighbour-list cutoff */
  float weight_force, weight_energy, weight_stress;
  float* pmin;         /* [num_variables] */
  float* pmax;
} orc_potential;

typedef struct {
  int Nc, Nc_force, N, N_force, max_Na;
  int* Na;          /* [Nc] */
  int* Na_sum;      /* [Nc] */
  int* type;        /* [N] */
  float* h;         /* [Nc*18]: box columns a,b,c then inverse */
  float* r;         /* [3N] SoA x|y|z */
  float* pe_ref;    /* [Nc] */
  float* virial_ref;/* [6*Nc] component-major xx,yy,zz,xy,yz,zx */
  float* force_ref; /* [3N] SoA */
  /* normalisers, FP64 (the reference accumulates them in float) */
  double energy_minimum;
  double force_square_sum, potential_square_sum, virial_square_sum;
} orc_trainset;

typedef struct {
  int N;
  int* NN;      /* [N] neighbour count */
  int* start;   /* [N] offset of atom i's list in NL */
  int* NL;      /* CSR, ascending j per atom */
  int total;
} orc_neighbor;

typedef struct {
  int maximum_generation, population_size, parent_number;
  float mutation_rate;
} orc_ga_params;

/* All return 0 on success, non-zero with a message in orc_last_error(). */
const char* orc_last_error(void);
int orc_read_potential(const char* dir, orc_potential* pot);
int orc_read_train(const char* dir, orc_trainset* ts);
int orc_read_ga(const char* dir, orc_ga_params* ga);
void orc_free_potential(orc_potential* pot);
void orc_free_train(orc_trainset* ts);
void orc_free_neighbor(orc_neighbor* nb);

/* Builds a training set from arrays already in memory (copies them); the
 * normalisers are computed as the reader computes them. */
int orc_make_train(int Nc, int Nc_force, const int* Na, const int* type, const float* box9,
                   const float* r, const float* pe_ref, const float* virial_ref,
                   const float* force_ref, orc_trainset* ts);

/* ------------------------------------------------ geometry (FP32 exact) */
void orc_box_inverse(float* h18);
void orc_mic_f32(const float* h18, float* x, float* y, float* z);
int orc_build_neighbor(const orc_trainset* ts, float cutoff, orc_neighbor* nb);

/* --------------------------------------------------- potentials (FP64)
 * params: physical parameters as the reference's std::vector<float>.
 * Outputs per atom: f[3N] SoA, virial[6N] component-major, pe[N]. */
void orc_tersoff(int num_types, const float* params, const orc_trainset* ts,
                 const orc_neighbor* nb, double* f, double* virial, double* pe);
void orc_ri(const float* params, const orc_trainset* ts, const orc_neighbor* nb,
            double* f, double* virial, double* pe);
void orc_nn2b(int num_neurons, int species, const float* params, const orc_trainset* ts,
              const orc_neighbor* nb, double* f, double* virial, double* pe);
int orc_find_force(const orc_potential* pot, const float* params, const orc_trainset* ts,
                   const orc_neighbor* nb, double* f, double* virial, double* pe);

/* ------------------------------------------------------- fitness (FP64)
 * terms[0..2] = relative RMSE of energy, virial, force;
 * returns w_E*terms[0] + w_W*terms[1] + w_F*terms[2]. */
double orc_error_terms(const orc_potential* pot, const orc_trainset* ts, const double* f,
                       const double* virial, const double* pe, double terms[3]);
/* genes in [0,1] -> physical parameters, float arithmetic as the reference */
void orc_denormalise(const orc_potential* pot, const float* genes, float* params);
/* fitness of P individuals; threads > 1 uses OpenMP over individuals.
 * terms may be NULL, else [P][3]. */
int orc_fitness(const orc_potential* pot, const orc_trainset* ts, const orc_neighbor* nb,
                int P, const float* genes, double* fitness, double* terms, int threads);

/* ------------------------------------------------------------ RNG + GA */
typedef struct {
  uint32_t mt[624];
  int idx;
} orc_mt19937;

void orc_mt_seed(orc_mt19937* g, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* g);
float orc_canonical_float(orc_mt19937* g);              /* U[0,1) float */
int orc_uniform_int(orc_mt19937* g, int a, int b);       /* U{a..b}      */

typedef struct {
  orc_ga_params p;
  int V;
  int child_number;
  float* population;   /* [P*V] */
  float* population_copy;
  float* fitness;      /* [P]   */
  int* index;          /* [P]   */
  float* cumulative_probabilities; /* [parent_number] */
  orc_mt19937 rng;
} orc_ga;

int orc_ga_init(orc_ga* ga, const orc_ga_params* p, int V, uint32_t seed);
void orc_ga_free(orc_ga* ga);
void orc_ga_sort(orc_ga* ga);                 /* sorts ga->fitness, permutes rows */
void orc_ga_crossover(orc_ga* ga);
void orc_ga_mutation(orc_ga* ga, int generation);
/* one generation after fitness has been written into ga->fitness */
void orc_ga_step(orc_ga* ga, int generation);

#ifdef __cplusplus
}
#endif
#endif
