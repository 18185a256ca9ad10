/*
 * FP64 minimal Tersoff (test infrastructure only, see oracle.h).
 * Restates src/minimal_tersoff.cu:
 *   parameters and derived constants   :39-60
 *   fR, fA, fc, g and derivatives      :62-119
 *   step 1, bond order b and b'        :122-188
 *   step 2, partial forces f_ij, E_i   :191-316
 *   step 3, F_i and per-atom virial    :318-405
 * One-species parameter sets (num_types == 1) use the single set for every
 * pair; the reference leaves the second set unset in that case
 * (update_potential :45-52), so mixed-type pairs are only defined for
 * num_types == 2.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double D0[2], A[2], R0[2], S[2], EN, BETA, H, R1, R2, PI_FACTOR, MINUS_HALF_OVER_N;
} tparams;

static void set_params(int num_types, const float* p, tparams* t)
{
  int off = 0;
  t->D0[0] = p[0]; t->A[0] = p[1]; t->R0[0] = p[2]; t->S[0] = p[3];
  if (num_types == 2) {
    off = 4;
    t->D0[1] = p[4]; t->A[1] = p[5]; t->R0[1] = p[6]; t->S[1] = p[7];
  } else {
    t->D0[1] = t->D0[0]; t->A[1] = t->A[0]; t->R0[1] = t->R0[0]; t->S[1] = t->S[0];
  }
  t->EN = p[4 + off];
  t->BETA = p[5 + off];
  t->H = p[6 + off];
  t->R1 = p[7 + off];
  t->R2 = p[8 + off];
  t->PI_FACTOR = M_PI / (t->R2 - t->R1);
  t->MINUS_HALF_OVER_N = -0.5 / t->EN;
}

static void fr_frp(double d0, double a, double r0, double s, double d, double* fr, double* frp)
{
  *fr = d0 / (s - 1.0) * exp(-sqrt(2.0 * s) * a * (d - r0));
  *frp = -sqrt(2.0 * s) * a * *fr;
}

static void fa_fap(double d0, double a, double r0, double s, double d, double* fa, double* fap)
{
  *fa = s * d0 / (s - 1.0) * exp(-sqrt(2.0 / s) * a * (d - r0));
  *fap = -sqrt(2.0 / s) * a * *fa;
}

static void fc_fcp(const tparams* t, double d, double* fc, double* fcp)
{
  if (d < t->R1) {
    *fc = 1.0;
    *fcp = 0.0;
  } else if (d < t->R2) {
    *fc = 0.5 * cos(t->PI_FACTOR * (d - t->R1)) + 0.5;
    *fcp = -sin(t->PI_FACTOR * (d - t->R1)) * t->PI_FACTOR * 0.5;
  } else {
    *fc = 0.0;
    *fcp = 0.0;
  }
}

/* MIC displacement in FP64 following src/mic.cuh:20-32 with the float box */
static void mic64(const float* h, double* x, double* y, double* z)
{
  double sx = (double)h[9] * *x + (double)h[10] * *y + (double)h[11] * *z;
  double sy = (double)h[12] * *x + (double)h[13] * *y + (double)h[14] * *z;
  double sz = (double)h[15] * *x + (double)h[16] * *y + (double)h[17] * *z;
  sx -= nearbyint(sx);
  sy -= nearbyint(sy);
  sz -= nearbyint(sz);
  *x = h[0] * sx + h[1] * sy + h[2] * sz;
  *y = h[3] * sx + h[4] * sy + h[5] * sz;
  *z = h[6] * sx + h[7] * sy + h[8] * sz;
}

static void disp(const orc_trainset* ts, const float* h, int i, int j, double* v)
{
  const int N = ts->N;
  v[0] = (double)ts->r[j] - (double)ts->r[i];
  v[1] = (double)ts->r[j + N] - (double)ts->r[i + N];
  v[2] = (double)ts->r[j + 2 * N] - (double)ts->r[i + 2 * N];
  mic64(h, &v[0], &v[1], &v[2]);
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

void orc_tersoff(int num_types, const float* params, const orc_trainset* ts,
                 const orc_neighbor* nb, double* f, double* virial, double* pe)
{
  tparams t;
  set_params(num_types, params, &t);
  const int N = ts->N;
  const int total = nb->total;
  double* b = (double*)calloc(total + 1, sizeof(double));
  double* bp = (double*)calloc(total + 1, sizeof(double));
  double* f12 = (double*)calloc(3 * (size_t)(total + 1), sizeof(double));

  /* step 1 (:122-188): zeta_ij, b_ij, b'_ij */
  for (int i = 0; i < N; ++i) {
    const float* h = ts->h + 18 * config_of(ts, i);
    const int s0 = nb->start[i], nn = nb->NN[i];
    for (int s = 0; s < nn; ++s) {
      int j = nb->NL[s0 + s];
      double r12[3];
      disp(ts, h, i, j, r12);
      double d12 = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
      double zeta = 0.0;
      for (int u = 0; u < nn; ++u) {
        int k = nb->NL[s0 + u];
        if (k == j) continue;
        double r13[3];
        disp(ts, h, i, k, r13);
        double d13 = sqrt(r13[0] * r13[0] + r13[1] * r13[1] + r13[2] * r13[2]);
        double cos123 = (r12[0] * r13[0] + r12[1] * r13[1] + r12[2] * r13[2]) / (d12 * d13);
        double fc13, fcp13;
        fc_fcp(&t, d13, &fc13, &fcp13);
        double x = cos123 - t.H;
        zeta += fc13 * t.BETA * x * x;
      }
      if (zeta < 1.0e-16) {
        b[s0 + s] = 1.0;
        bp[s0 + s] = 0.0;
      } else {
        double bzn = pow(zeta, t.EN);
        double bij = pow(1.0 + bzn, t.MINUS_HALF_OVER_N);
        b[s0 + s] = bij;
        bp[s0 + s] = -bij * bzn * 0.5 / ((1.0 + bzn) * zeta);
      }
    }
  }

  /* step 2 (:191-316): partial forces and site energies */
  for (int i = 0; i < N; ++i) {
    const float* h = ts->h + 18 * config_of(ts, i);
    const int s0 = nb->start[i], nn = nb->NN[i];
    const int ti = ts->type[i];
    double e = 0.0;
    for (int s = 0; s < nn; ++s) {
      int j = nb->NL[s0 + s];
      int p12 = (ti != ts->type[j]) ? 1 : 0;
      double r12[3];
      disp(ts, h, i, j, r12);
      double d12 = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
      double d12inv = 1.0 / d12;
      double fc12, fcp12, fa12, fap12, fr12, frp12;
      fc_fcp(&t, d12, &fc12, &fcp12);
      fa_fap(t.D0[p12], t.A[p12], t.R0[p12], t.S[p12], d12, &fa12, &fap12);
      fr_frp(t.D0[p12], t.A[p12], t.R0[p12], t.S[p12], d12, &fr12, &frp12);
      double b12 = b[s0 + s];
      double factor3 = (fcp12 * (fr12 - b12 * fa12) + fc12 * (frp12 - b12 * fap12)) * d12inv;
      double fx = r12[0] * factor3 * 0.5, fy = r12[1] * factor3 * 0.5, fz = r12[2] * factor3 * 0.5;
      e += fc12 * (fr12 - b12 * fa12) * 0.5;
      double bp12 = bp[s0 + s];
      for (int u = 0; u < nn; ++u) {
        int k = nb->NL[s0 + u];
        if (k == j) continue;
        int p13 = (ti != ts->type[k]) ? 1 : 0;
        double r13[3];
        disp(ts, h, i, k, r13);
        double d13 = sqrt(r13[0] * r13[0] + r13[1] * r13[1] + r13[2] * r13[2]);
        double fc13, fcp13, fa13, fap13;
        fc_fcp(&t, d13, &fc13, &fcp13);
        fa_fap(t.D0[p13], t.A[p13], t.R0[p13], t.S[p13], d13, &fa13, &fap13);
        double bp13 = bp[s0 + u];
        double one_over_d12d13 = 1.0 / (d12 * d13);
        double cos123 = (r12[0] * r13[0] + r12[1] * r13[1] + r12[2] * r13[2]) * one_over_d12d13;
        double cos123_over_d12d12 = cos123 * d12inv * d12inv;
        double x = cos123 - t.H;
        double g123 = t.BETA * x * x, gp123 = t.BETA * 2.0 * x;
        double dc = -fc12 * bp12 * fa12 * fc13 * gp123 - fc12 * bp13 * fa13 * fc13 * gp123;
        double dr = -fcp12 * bp13 * fa13 * g123 * fc13 * d12inv;
        fx += (r12[0] * dr + dc * (r13[0] * one_over_d12d13 - r12[0] * cos123_over_d12d12)) * 0.5;
        fy += (r12[1] * dr + dc * (r13[1] * one_over_d12d13 - r12[1] * cos123_over_d12d12)) * 0.5;
        fz += (r12[2] * dr + dc * (r13[2] * one_over_d12d13 - r12[2] * cos123_over_d12d12)) * 0.5;
      }
      f12[3 * (s0 + s) + 0] = fx;
      f12[3 * (s0 + s) + 1] = fy;
      f12[3 * (s0 + s) + 2] = fz;
    }
    pe[i] = e;
  }

  /* step 3 (:318-405): F_i = sum_j (f_ij - f_ji); unsymmetrised virial */
  for (int i = 0; i < N; ++i) {
    const float* h = ts->h + 18 * config_of(ts, i);
    const int s0 = nb->start[i], nn = nb->NN[i];
    double F[3] = {0, 0, 0}, W[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < nn; ++s) {
      int j = nb->NL[s0 + s];
      double r12[3];
      disp(ts, h, i, j, r12);
      int off = 0; /* slot of i in j's list; 0 if absent (:368-375) */
      for (int u = 0; u < nb->NN[j]; ++u)
        if (nb->NL[nb->start[j] + u] == i) {
          off = u;
          break;
        }
      const double* a = f12 + 3 * (s0 + s);
      const double* r = f12 + 3 * (nb->start[j] + off);
      double dx = a[0] - r[0], dy = a[1] - r[1], dz = a[2] - r[2];
      F[0] += dx; F[1] += dy; F[2] += dz;
      W[0] -= r12[0] * dx * 0.5;
      W[1] -= r12[1] * dy * 0.5;
      W[2] -= r12[2] * dz * 0.5;
      W[3] -= r12[0] * dy * 0.5;
      W[4] -= r12[1] * dz * 0.5;
      W[5] -= r12[2] * dx * 0.5;
    }
    for (int d = 0; d < 3; ++d) f[i + N * d] = F[d];
    for (int k = 0; k < 6; ++k) virial[i + N * k] = W[k];
  }
  free(b);
  free(bp);
  free(f12);
}
