/*
 * FP64 fitness reduction (test infrastructure only, see oracle.h).
 *   denormalisation      src/fitness.cu:319-324 (float arithmetic, no FMA)
 *   energy term          src/fitness.cu:446-504
 *   virial term          src/fitness.cu:506-555
 *   force term           src/fitness.cu:392-444
 *   weighted sum         src/fitness.cu:329-331
 * Energies are per-configuration totals; E_ref is not shifted in the
 * numerator, E_min only enters the normaliser (src/fitness.cu:154-160).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_denormalise(const orc_potential* pot, const float* genes, float* params)
{
  for (int m = 0; m < pot->num_variables; ++m) {
    float a = pot->pmin[m];
    float b = pot->pmax[m] - a;
    float bx = b * genes[m];
    params[m] = a + bx;
  }
}

double orc_error_terms(const orc_potential* pot, const orc_trainset* ts, const double* f,
                       const double* virial, const double* pe, double terms[3])
{
  const int N = ts->N, Nc = ts->Nc;
  double se = 0.0, sv = 0.0, sf = 0.0;
  for (int c = ts->Nc_force; c < Nc; ++c) {
    const int b = ts->Na_sum[c], e = b + ts->Na[c];
    double E = 0.0;
    for (int i = b; i < e; ++i) E += pe[i];
    double d = E - (double)ts->pe_ref[c];
    se += d * d;
    for (int k = 0; k < 6; ++k) {
      double W = 0.0;
      for (int i = b; i < e; ++i) W += virial[i + N * k];
      double dv = W - (double)ts->virial_ref[c + Nc * k];
      sv += dv * dv;
    }
  }
  for (int i = 0; i < ts->N_force; ++i)
    for (int k = 0; k < 3; ++k) {
      double d = f[i + N * k] - (double)ts->force_ref[i + N * k];
      sf += d * d;
    }
  terms[0] = sqrt(se / ts->potential_square_sum);
  terms[1] = sqrt(sv / ts->virial_square_sum);
  terms[2] = sqrt(sf / ts->force_square_sum);
  return (double)pot->weight_energy * terms[0] + (double)pot->weight_stress * terms[1] +
         (double)pot->weight_force * terms[2];
}

int orc_fitness(const orc_potential* pot, const orc_trainset* ts, const orc_neighbor* nb, int P,
                const float* genes, double* fitness, double* terms, int threads)
{
  const int N = ts->N, V = pot->num_variables;
  int rc = 0;
#ifdef _OPENMP
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(| : rc)
#endif
  for (int p = 0; p < P; ++p) {
    double* f = (double*)malloc(sizeof(double) * 3 * N);
    double* w = (double*)malloc(sizeof(double) * 6 * N);
    double* e = (double*)malloc(sizeof(double) * N);
    float* params = (float*)malloc(sizeof(float) * V);
    orc_denormalise(pot, genes + (size_t)p * V, params);
    rc |= orc_find_force(pot, params, ts, nb, f, w, e);
    double t[3];
    fitness[p] = orc_error_terms(pot, ts, f, w, e, t);
    if (terms) {
      terms[3 * p + 0] = t[0];
      terms[3 * p + 1] = t[1];
      terms[3 * p + 2] = t[2];
    }
    free(f);
    free(w);
    free(e);
    free(params);
  }
  (void)threads;
  return rc;
}
