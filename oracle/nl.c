/*
 * Oracle geometry (test infrastructure only, see oracle.h).
 *
 * FP32 minimum-image convention and neighbour list exactly as the source
 * expressions of the reference are written (left-to-right, every product and
 * sum rounded to float, nearbyint = round-half-even), compiled with
 * -ffp-contract=off so no fused multiply-add changes a rounding.
 *   MIC  src/mic.cuh:20-32
 *   NL   src/neighbor.cu:30-66 (j ascending within the configuration, j != i,
 *        strict d^2 < rc^2, rc^2 formed in float as neighbor.cu:72)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

void orc_mic_f32(const float* h, float* x, float* y, float* z)
{
  float sx = h[9] * *x + h[10] * *y + h[11] * *z;
  float sy = h[12] * *x + h[13] * *y + h[14] * *z;
  float sz = h[15] * *x + h[16] * *y + h[17] * *z;
  sx -= nearbyintf(sx);
  sy -= nearbyintf(sy);
  sz -= nearbyintf(sz);
  *x = h[0] * sx + h[1] * sy + h[2] * sz;
  *y = h[3] * sx + h[4] * sy + h[5] * sz;
  *z = h[6] * sx + h[7] * sy + h[8] * sz;
}

static int is_neighbor(const orc_trainset* ts, const float* h, int i, int j, float rc2)
{
  const int N = ts->N;
  float x = ts->r[j] - ts->r[i];
  float y = ts->r[j + N] - ts->r[i + N];
  float z = ts->r[j + 2 * N] - ts->r[i + 2 * N];
  orc_mic_f32(h, &x, &y, &z);
  float d2 = x * x + y * y + z * z;
  return d2 < rc2;
}

int orc_build_neighbor(const orc_trainset* ts, float cutoff, orc_neighbor* nb)
{
  const int N = ts->N;
  const float rc2 = cutoff * cutoff;
  nb->N = N;
  nb->NN = (int*)calloc(N, sizeof(int));
  nb->start = (int*)calloc(N, sizeof(int));
  long total = 0;
  for (int c = 0; c < ts->Nc; ++c) {
    const float* h = ts->h + 18 * c;
    const int b = ts->Na_sum[c], e = b + ts->Na[c];
    for (int i = b; i < e; ++i) {
      int cnt = 0;
      for (int j = b; j < e; ++j)
        if (j != i && is_neighbor(ts, h, i, j, rc2)) ++cnt;
      nb->NN[i] = cnt;
      nb->start[i] = (int)total;
      total += cnt;
    }
  }
  nb->total = (int)total;
  nb->NL = (int*)malloc(sizeof(int) * (total > 0 ? total : 1));
  for (int c = 0; c < ts->Nc; ++c) {
    const float* h = ts->h + 18 * c;
    const int b = ts->Na_sum[c], e = b + ts->Na[c];
    for (int i = b; i < e; ++i) {
      int k = nb->start[i];
      for (int j = b; j < e; ++j)
        if (j != i && is_neighbor(ts, h, i, j, rc2)) nb->NL[k++] = j;
    }
  }
  return 0;
}

void orc_free_neighbor(orc_neighbor* nb)
{
  free(nb->NN);
  free(nb->start);
  free(nb->NL);
  nb->NN = nb->start = nb->NL = NULL;
}
