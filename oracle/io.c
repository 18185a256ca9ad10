/*
 * Oracle input readers (test infrastructure only, see oracle.h).
 * Restates the reference's parsers:
 *   potential.in  src/fitness.cu:220-314
 *   train.in      src/fitness.cu:59-161 (read_Nc :163-182, read_Na :184-218)
 *   ga.in         src/ga.cu:214-316 with the tokenizer src/read_file.cu:46-102
 */
#include "oracle.h"

#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];

const char* orc_last_error(void) { return g_err; }

static int fail(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return 1;
}

static FILE* open_in(const char* dir, const char* name)
{
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  return fopen(path, "r");
}

/* src/fitness.cu:220-314: fixed order, keyword tokens ignored */
int orc_read_potential(const char* dir, orc_potential* pot)
{
  memset(pot, 0, sizeof *pot);
  FILE* fid = open_in(dir, "potential.in");
  if (!fid) return fail("Failed to open %s/potential.in", dir);
  char name[256];
  int rc = 0;
  if (fscanf(fid, "%255s%d", name, &pot->potential_type) != 2) {
    rc = fail("reading error for potential.in.");
    goto done;
  }
  switch (pot->potential_type) {
  case 0: pot->num_variables = 15; break;
  case 1: pot->num_variables = 9; break;
  case 2: pot->num_variables = 13; break;
  case 3:
  case 4: /* 4: the species-pair input, n more weights */
    if (fscanf(fid, "%255s%d", name, &pot->num_neurons) != 2) {
      rc = fail("reading error for potential.in.");
      goto done;
    }
    pot->num_variables = pot->num_neurons * (pot->num_neurons + 4) + 1 +
                         (pot->potential_type == 4 ? pot->num_neurons : 0);
    break;
  default: rc = fail("unsupported potential type."); goto done;
  }
  pot->pmin = (float*)calloc(pot->num_variables, sizeof(float));
  pot->pmax = (float*)calloc(pot->num_variables, sizeof(float));
  if (fscanf(fid, "%255s%f", name, &pot->cutoff) != 2 ||
      fscanf(fid, "%255s%f", name, &pot->weight_force) != 2 ||
      fscanf(fid, "%255s%f", name, &pot->weight_energy) != 2 ||
      fscanf(fid, "%255s%f", name, &pot->weight_stress) != 2) {
    rc = fail("reading error for potential.in.");
    goto done;
  }
  if (pot->weight_force < 0 || pot->weight_energy < 0 || pot->weight_stress < 0) {
    rc = fail("weights should >= 0");
    goto done;
  }
  for (int n = 0; n < pot->num_variables; ++n) {
    if (pot->potential_type < 3) {
      if (fscanf(fid, "%255s%f%f", name, &pot->pmin[n], &pot->pmax[n]) != 3) {
        rc = fail("reading error for potential.in.");
        goto done;
      }
    } else {
      pot->pmin[n] = -3.0f;
      pot->pmax[n] = 3.0f;
    }
  }
done:
  fclose(fid);
  return rc;
}

void orc_free_potential(orc_potential* pot)
{
  free(pot->pmin);
  free(pot->pmax);
  pot->pmin = pot->pmax = NULL;
}

/* src/fitness.cu:40-57, FP32 exactly as written (no contraction: this file
 * is compiled with -ffp-contract=off) */
void orc_box_inverse(float* h)
{
  h[9] = h[4] * h[8] - h[5] * h[7];
  h[10] = h[2] * h[7] - h[1] * h[8];
  h[11] = h[1] * h[5] - h[2] * h[4];
  h[12] = h[5] * h[6] - h[3] * h[8];
  h[13] = h[0] * h[8] - h[2] * h[6];
  h[14] = h[2] * h[3] - h[0] * h[5];
  h[15] = h[3] * h[7] - h[4] * h[6];
  h[16] = h[1] * h[6] - h[0] * h[7];
  h[17] = h[0] * h[4] - h[1] * h[3];
  float volume = h[0] * (h[4] * h[8] - h[5] * h[7]) + h[1] * (h[5] * h[6] - h[3] * h[8]) +
                 h[2] * (h[3] * h[7] - h[4] * h[6]);
  for (int n = 9; n < 18; n++) h[n] /= volume;
}

static void compute_normalisers(orc_trainset* ts)
{
  /* src/fitness.cu:86-89,103-104,138-140,154-160 (accumulated in FP64 here) */
  double emin = 0.0;
  for (int n = ts->Nc_force; n < ts->Nc; ++n)
    if (ts->pe_ref[n] < emin) emin = ts->pe_ref[n];
  ts->energy_minimum = emin;
  double fs = 0, ps = 0, vs = 0;
  for (int i = 0; i < ts->N_force; ++i)
    for (int d = 0; d < 3; ++d) {
      double v = ts->force_ref[i + ts->N * d];
      fs += v * v;
    }
  for (int n = ts->Nc_force; n < ts->Nc; ++n) {
    double e = (double)ts->pe_ref[n] - emin;
    ps += e * e;
    for (int k = 0; k < 6; ++k) {
      double v = ts->virial_ref[n + ts->Nc * k];
      vs += v * v;
    }
  }
  ts->force_square_sum = fs;
  ts->potential_square_sum = ps;
  ts->virial_square_sum = vs;
}

static int alloc_train(orc_trainset* ts)
{
  ts->h = (float*)calloc((size_t)ts->Nc * 18, sizeof(float));
  ts->pe_ref = (float*)calloc(ts->Nc, sizeof(float));
  ts->virial_ref = (float*)calloc((size_t)ts->Nc * 6, sizeof(float));
  ts->type = (int*)calloc(ts->N, sizeof(int));
  ts->r = (float*)calloc((size_t)ts->N * 3, sizeof(float));
  ts->force_ref = (float*)calloc((size_t)ts->N * 3, sizeof(float));
  return !(ts->h && ts->pe_ref && ts->virial_ref && ts->type && ts->r && ts->force_ref);
}

static int finish_counts(orc_trainset* ts)
{
  ts->N = 0;
  ts->max_Na = 0;
  for (int c = 0; c < ts->Nc; ++c) {
    if (ts->Na[c] < 2) return fail("Number of atoms %d should >= 2", ts->Na[c]);
    ts->Na_sum[c] = ts->N;
    ts->N += ts->Na[c];
    if (ts->Na[c] > ts->max_Na) ts->max_Na = ts->Na[c];
  }
  ts->N_force = 0;
  for (int c = 0; c < ts->Nc_force; ++c) ts->N_force += ts->Na[c];
  return 0;
}

static int check_nc(int Nc, int Nc_force)
{
  /* src/fitness.cu:169-177 */
  if (Nc < 2) return fail("Number of configurations should >= 2");
  if (Nc_force < 1) return fail("Number of force configurations should >= 1");
  if (Nc_force > Nc - 1) return fail("Number of potential configurations should >= 1");
  return 0;
}

int orc_read_train(const char* dir, orc_trainset* ts)
{
  memset(ts, 0, sizeof *ts);
  FILE* fid = open_in(dir, "train.in");
  if (!fid) return fail("Failed to open %s/train.in", dir);
  int rc = 0;
  if (fscanf(fid, "%d%d", &ts->Nc, &ts->Nc_force) != 2) {
    rc = fail("Reading error for train.in.");
    goto done;
  }
  if ((rc = check_nc(ts->Nc, ts->Nc_force))) goto done;
  ts->Na = (int*)calloc(ts->Nc, sizeof(int));
  ts->Na_sum = (int*)calloc(ts->Nc, sizeof(int));
  for (int c = 0; c < ts->Nc; ++c)
    if (fscanf(fid, "%d", &ts->Na[c]) != 1) {
      rc = fail("Reading error for train.in.");
      goto done;
    }
  if ((rc = finish_counts(ts))) goto done;
  if (alloc_train(ts)) {
    rc = fail("out of memory");
    goto done;
  }
  const int N = ts->N, Nc = ts->Nc;
  for (int n = 0; n < Nc; ++n) {
    if (n >= ts->Nc_force) {
      float* v = ts->virial_ref;
      if (fscanf(fid, "%f%f%f%f%f%f%f", &ts->pe_ref[n], &v[n], &v[n + Nc], &v[n + 2 * Nc],
                 &v[n + 3 * Nc], &v[n + 4 * Nc], &v[n + 5 * Nc]) != 7) {
        rc = fail("reading error for train.in.");
        goto done;
      }
    }
    float t[9];
    for (int k = 0; k < 9; ++k)
      if (fscanf(fid, "%f", &t[k]) != 1) {
        rc = fail("reading error for train.in.");
        goto done;
      }
    /* box stored as columns a, b, c (transpose of the file's rows) */
    float* h = ts->h + 18 * n;
    h[0] = t[0]; h[3] = t[1]; h[6] = t[2];
    h[1] = t[3]; h[4] = t[4]; h[7] = t[5];
    h[2] = t[6]; h[5] = t[7]; h[8] = t[8];
    orc_box_inverse(h);
    for (int k = 0; k < ts->Na[n]; ++k) {
      int i = ts->Na_sum[n] + k;
      if (n < ts->Nc_force) {
        if (fscanf(fid, "%d%f%f%f%f%f%f", &ts->type[i], &ts->r[i], &ts->r[i + N],
                   &ts->r[i + 2 * N], &ts->force_ref[i], &ts->force_ref[i + N],
                   &ts->force_ref[i + 2 * N]) != 7) {
          rc = fail("reading error for train.in.");
          goto done;
        }
      } else {
        if (fscanf(fid, "%d%f%f%f", &ts->type[i], &ts->r[i], &ts->r[i + N], &ts->r[i + 2 * N]) !=
            4) {
          rc = fail("reading error for train.in.");
          goto done;
        }
      }
    }
  }
  compute_normalisers(ts);
done:
  fclose(fid);
  return rc;
}

int orc_make_train(int Nc, int Nc_force, const int* Na, const int* type, const float* box9,
                   const float* r, const float* pe_ref, const float* virial_ref,
                   const float* force_ref, orc_trainset* ts)
{
  memset(ts, 0, sizeof *ts);
  if (check_nc(Nc, Nc_force)) return 1;
  ts->Nc = Nc;
  ts->Nc_force = Nc_force;
  ts->Na = (int*)calloc(Nc, sizeof(int));
  ts->Na_sum = (int*)calloc(Nc, sizeof(int));
  memcpy(ts->Na, Na, sizeof(int) * Nc);
  if (finish_counts(ts)) return 1;
  if (alloc_train(ts)) return fail("out of memory");
  memcpy(ts->type, type, sizeof(int) * ts->N);
  memcpy(ts->r, r, sizeof(float) * 3 * ts->N);
  if (force_ref) memcpy(ts->force_ref, force_ref, sizeof(float) * 3 * ts->N);
  if (pe_ref) memcpy(ts->pe_ref, pe_ref, sizeof(float) * Nc);
  if (virial_ref) memcpy(ts->virial_ref, virial_ref, sizeof(float) * 6 * Nc);
  for (int n = 0; n < Nc; ++n) {
    const float* t = box9 + 9 * n;
    float* h = ts->h + 18 * n;
    h[0] = t[0]; h[3] = t[1]; h[6] = t[2];
    h[1] = t[3]; h[4] = t[4]; h[7] = t[5];
    h[2] = t[6]; h[5] = t[7]; h[8] = t[8];
    orc_box_inverse(h);
  }
  compute_normalisers(ts);
  return 0;
}

void orc_free_train(orc_trainset* ts)
{
  free(ts->Na); free(ts->Na_sum); free(ts->type); free(ts->h); free(ts->r);
  free(ts->pe_ref); free(ts->virial_ref); free(ts->force_ref);
  memset(ts, 0, sizeof *ts);
}

/* ---- ga.in: src/read_file.cu:46-102 tokenizer, src/ga.cu:214-316 keys ---- */

static int valid_int(const char* s, int* out)
{
  if (!s || !*s) return 0;
  char* p;
  errno = 0;
  long v = strtol(s, &p, 0);
  if (errno || p == s || *p) return 0;
  *out = (int)v;
  return 1;
}

static int valid_real(const char* s, float* out)
{
  if (!s || !*s) return 0;
  char* p;
  errno = 0;
  double v = strtod(s, &p);
  if (errno || p == s || *p) return 0;
  *out = (float)v;
  return 1;
}

int orc_read_ga(const char* dir, orc_ga_params* ga)
{
  ga->maximum_generation = 1000;
  ga->population_size = 200;
  ga->parent_number = 100;
  ga->mutation_rate = 0.2f;
  FILE* fid = open_in(dir, "ga.in");
  if (!fid) return fail("Failed to open %s/ga.in", dir);
  char line[4096];
  int rc = 0;
  while (fgets(line, sizeof line, fid)) {
    char* hash = strchr(line, '#');
    if (hash) *hash = 0;
    char* tok[16];
    int nt = 0;
    for (char* t = strtok(line, " \t\r\n\v\f"); t && nt < 16; t = strtok(NULL, " \t\r\n\v\f"))
      tok[nt++] = t;
    if (nt == 0) continue;
    if (nt != 2) {
      rc = fail("%s should have 1 parameter.", tok[0]);
      break;
    }
    if (!strcmp(tok[0], "maximum_generation")) {
      if (!valid_int(tok[1], &ga->maximum_generation) || ga->maximum_generation < 1) {
        rc = fail("maximum_generation should be a positive integer.");
        break;
      }
    } else if (!strcmp(tok[0], "population_size")) {
      if (!valid_int(tok[1], &ga->population_size) || ga->population_size < 20) {
        rc = fail("population_size should be an integer >= 20.");
        break;
      }
    } else if (!strcmp(tok[0], "parent_number")) {
      if (!valid_int(tok[1], &ga->parent_number) || ga->parent_number < 10) {
        rc = fail("parent_number should be an integer >= 10.");
        break;
      }
    } else if (!strcmp(tok[0], "mutation_rate")) {
      if (!valid_real(tok[1], &ga->mutation_rate) || ga->mutation_rate < 0 ||
          ga->mutation_rate > 1) {
        rc = fail("mutation_rate should be in [0, 1].");
        break;
      }
    } else {
      rc = fail("'%s' is invalid keyword.", tok[0]);
      break;
    }
  }
  fclose(fid);
  return rc;
}
