/*
 * FP64 pair potentials (test infrastructure only, see oracle.h).
 *   RI    src/ri.cu:26-74 (constants, update_potential), :77-110 (U, U'/r),
 *         :112-187 (pair loop)
 *   NN2B  src/nn2b.cu:33-46 (parameter layout), :49-78 (MLP and derivative),
 *         :80-92 (cutoff), :94-171 (pair loop, including f2 = (p'fc + U fc')/r
 *         with U = p fc, nn2b.cu:143-148), src/nn2b.cuh:36-38 (r1 = 7,
 *         r2 = 8, pi_factor = 3.1415927f, scaling = 1)
 *   NN2B with a species-pair input (potential type 4; no counterpart in the
 *         reference): the same network and pair loop, with the first layer
 *         x1 = tanh(w0 r + ws s - b0), s = t_i + t_j, and the n weights ws
 *         after the reference's n (n + 4) + 1 parameters.  Pinned by: ws = 0
 *         gives type 3 exactly; finite differences; a dimer closed form.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

#define RI_ALPHA 0.2
#define RI_ALPHA_SQ 0.04
#define RI_PI_FACTOR 0.225675833419103 /* 2 alpha / sqrt(pi) */
#define K_C 14.399645

typedef struct {
  double a[3], b[3], c[3], qq[3], v_rc, dv_rc, cutoff, d0, alpha, r0, s;
} ripar;

static void ri_params(const float* p, ripar* q)
{
  double q0 = p[0];
  q->qq[0] = q0 * q0 * K_C;
  q->qq[1] = -0.5 * q0 * q0 * K_C;
  q->qq[2] = 0.25 * q0 * q0 * K_C;
  for (int t = 0; t < 3; ++t) {
    q->a[t] = p[1 + 3 * t];
    q->b[t] = 1.0 / (double)p[2 + 3 * t];
    q->c[t] = p[3 + 3 * t];
  }
  q->cutoff = p[10];
  q->d0 = p[11];
  q->alpha = p[12];
  q->r0 = p[13];
  q->s = p[14];
  double rc = q->cutoff;
  q->v_rc = erfc(RI_ALPHA * rc) / rc;
  q->dv_rc = -erfc(RI_ALPHA * rc) / (rc * rc) - RI_PI_FACTOR * exp(-RI_ALPHA_SQ * rc * rc) / rc;
}

static void ri_p2_f2(const ripar* q, int t, double d, double* p2, double* f2)
{
  double dinv = 1.0 / d;
  double dinv6 = pow(dinv, 6), dinv7 = dinv6 * dinv;
  double ex = exp(-d * q->b[t]);
  double erfc_r = erfc(RI_ALPHA * d) * dinv;
  double p = q->a[t] * ex - q->c[t] * dinv6;
  double g = 6.0 * q->c[t] * dinv7 - q->a[t] * ex * q->b[t];
  p += q->qq[t] * (erfc_r - q->v_rc - q->dv_rc * (d - q->cutoff));
  g -= q->qq[t] * (erfc_r * dinv + RI_PI_FACTOR * dinv * exp(-RI_ALPHA_SQ * d * d) + q->dv_rc);
  if (t == 1) {
    double fr = q->d0 / (q->s - 1.0) * exp(-sqrt(2.0 * q->s) * q->alpha * (d - q->r0));
    double frp = -sqrt(2.0 * q->s) * q->alpha * fr;
    double fa = q->s * q->d0 / (q->s - 1.0) * exp(-sqrt(2.0 / q->s) * q->alpha * (d - q->r0));
    double fap = -sqrt(2.0 / q->s) * q->alpha * fa;
    p += fr - fa;
    g += frp - fap;
  }
  *p2 = p;
  *f2 = g * dinv;
}

typedef struct {
  int n;
  const float* w0;
  const float* b0;
  const float* w1;
  const float* b1;
  const float* w2;
  const float* ws; /* species-pair weights (type 4), else NULL */
  double b2;
  double* x1;
  double* y1;
} nnpar;

/* s: the pair's species code t_i + t_j (used when q->ws is set) */
static void nn_p2_f2(const nnpar* q, double d, int s, double* p2, double* f2)
{
  const int n = q->n;
  for (int m = 0; m < n; ++m) {
    const double in = q->ws ? (double)q->ws[m] * s : 0.0;
    q->x1[m] = tanh((double)q->w0[m] * d + in - (double)q->b0[m]);
    q->y1[m] = (1.0 - q->x1[m] * q->x1[m]) * (double)q->w0[m];
  }
  double p = 0.0, g = 0.0;
  for (int k = 0; k < n; ++k) {
    double z = 0.0, y = 0.0;
    for (int m = 0; m < n; ++m) {
      z += (double)q->w1[k * n + m] * q->x1[m];
      y += (double)q->w1[k * n + m] * q->y1[m];
    }
    double x2 = tanh(z - (double)q->b1[k]);
    y *= 1.0 - x2 * x2;
    p += (double)q->w2[k] * x2;
    g += (double)q->w2[k] * y;
  }
  *p2 = p - q->b2;
  *f2 = g;
}

static int config_of(const orc_trainset* ts, int i)
{
  int lo = 0, hi = ts->Nc - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) / 2;
    if (ts->Na_sum[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

/* pair loop of src/ri.cu:133-187 and src/nn2b.cu:115-171 */
static void pair_loop(int kind, const void* par, const orc_trainset* ts, const orc_neighbor* nb,
                      double* f, double* virial, double* pe)
{
  const int N = ts->N;
  const double nn_r1 = 7.0, nn_r2 = 8.0, nn_pi = (double)3.1415927f;
  for (int i = 0; i < N; ++i) {
    const float* h = ts->h + 18 * config_of(ts, i);
    double F[3] = {0, 0, 0}, W[6] = {0, 0, 0, 0, 0, 0}, e = 0.0;
    for (int s = 0; s < nb->NN[i]; ++s) {
      int j = nb->NL[nb->start[i] + s];
      double x = (double)ts->r[j] - (double)ts->r[i];
      double y = (double)ts->r[j + N] - (double)ts->r[i + N];
      double z = (double)ts->r[j + 2 * N] - (double)ts->r[i + 2 * N];
      double sx = (double)h[9] * x + (double)h[10] * y + (double)h[11] * z;
      double sy = (double)h[12] * x + (double)h[13] * y + (double)h[14] * z;
      double sz = (double)h[15] * x + (double)h[16] * y + (double)h[17] * z;
      sx -= nearbyint(sx);
      sy -= nearbyint(sy);
      sz -= nearbyint(sz);
      x = h[0] * sx + h[1] * sy + h[2] * sz;
      y = h[3] * sx + h[4] * sy + h[5] * sz;
      z = h[6] * sx + h[7] * sy + h[8] * sz;
      double d = sqrt(x * x + y * y + z * z);
      double p2, f2;
      if (kind == 0) {
        ri_p2_f2((const ripar*)par, ts->type[i] + ts->type[j], d, &p2, &f2);
      } else {
        nn_p2_f2((const nnpar*)par, d, ts->type[i] + ts->type[j], &p2, &f2);
        double fc, fcp;
        if (d < nn_r1) {
          fc = 1.0;
          fcp = 0.0;
        } else if (d < nn_r2) {
          fc = 0.5 * cos(nn_pi * (d - nn_r1)) + 0.5;
          fcp = -sin(nn_pi * (d - nn_r1)) * nn_pi * 0.5;
        } else {
          fc = 0.0;
          fcp = 0.0;
        }
        p2 *= fc;
        f2 = (f2 * fc + p2 * fcp) / d;
      }
      F[0] += x * f2;
      F[1] += y * f2;
      F[2] += z * f2;
      W[0] -= x * x * f2 * 0.5;
      W[1] -= y * y * f2 * 0.5;
      W[2] -= z * z * f2 * 0.5;
      W[3] -= x * y * f2 * 0.5;
      W[4] -= y * z * f2 * 0.5;
      W[5] -= z * x * f2 * 0.5;
      e += p2 * 0.5;
    }
    for (int d = 0; d < 3; ++d) f[i + N * d] = F[d];
    for (int k = 0; k < 6; ++k) virial[i + N * k] = W[k];
    pe[i] = e;
  }
}

void orc_ri(const float* params, const orc_trainset* ts, const orc_neighbor* nb, double* f,
            double* virial, double* pe)
{
  ripar q;
  ri_params(params, &q);
  pair_loop(0, &q, ts, nb, f, virial, pe);
}

void orc_nn2b(int n, int species, const float* params, const orc_trainset* ts,
              const orc_neighbor* nb, double* f, double* virial, double* pe)
{
  nnpar q;
  q.n = n;
  q.ws = species ? params + n * (n + 4) + 1 : NULL;
  q.w0 = params;
  q.b0 = params + n;
  q.w1 = params + 2 * n;
  q.b1 = params + n * (n + 2);
  q.w2 = params + n * (n + 3);
  q.b2 = params[n * (n + 4)];
  q.x1 = (double*)malloc(sizeof(double) * n);
  q.y1 = (double*)malloc(sizeof(double) * n);
  pair_loop(1, &q, ts, nb, f, virial, pe);
  free(q.x1);
  free(q.y1);
}

int orc_find_force(const orc_potential* pot, const float* params, const orc_trainset* ts,
                   const orc_neighbor* nb, double* f, double* virial, double* pe)
{
  switch (pot->potential_type) {
  case 0: orc_ri(params, ts, nb, f, virial, pe); return 0;
  case 1: orc_tersoff(1, params, ts, nb, f, virial, pe); return 0;
  case 2: orc_tersoff(2, params, ts, nb, f, virial, pe); return 0;
  case 3: orc_nn2b(pot->num_neurons, 0, params, ts, nb, f, virial, pe); return 0;
  case 4: orc_nn2b(pot->num_neurons, 1, params, ts, nb, f, virial, pe); return 0;
  }
  return 1;
}
