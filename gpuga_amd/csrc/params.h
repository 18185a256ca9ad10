// Per-parameter-set constants of the three potential families, shared by
// host and device code.  Each struct is the by-value kernel argument of the
// single-parameter-set kernels and the per-lane register set of the
// population kernels; the derivation follows each reference update_potential.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace gpuga {

// Minimal Tersoff, src/minimal_tersoff.cu:39-60 and the prefactors of
// find_fr_and_frp / find_fa_and_fap (:62-75):
//   fR = cR * exp(kR * (d - r0)),  fA = cA * exp(kA * (d - r0))
// Set [1] is used for pairs of unlike types (:235-244); with one species it
// repeats set [0].
struct TersoffParams {
  float cR[2], kR[2], cA[2], kA[2], r0[2];
  float en, minus_half_over_n, beta, h, r1, r2, pi_factor;
};

__host__ __device__ inline TersoffParams tersoff_params(int num_types, const float* p)
{
  TersoffParams t;
  float D0[2], A[2], R0[2], S[2];
  D0[0] = p[0]; A[0] = p[1]; R0[0] = p[2]; S[0] = p[3];
  int off = 0;
  if (num_types == 2) {
    off = 4;
    D0[1] = p[4]; A[1] = p[5]; R0[1] = p[6]; S[1] = p[7];
  } else {
    D0[1] = D0[0]; A[1] = A[0]; R0[1] = R0[0]; S[1] = S[0];
  }
  for (int q = 0; q < 2; ++q) {
    t.cR[q] = D0[q] / (S[q] - 1.0f);
    t.kR[q] = -sqrtf(2.0f * S[q]) * A[q];
    t.cA[q] = S[q] * D0[q] / (S[q] - 1.0f);
    t.kA[q] = -sqrtf(2.0f / S[q]) * A[q];
    t.r0[q] = R0[q];
  }
  t.en = p[4 + off];
  t.beta = p[5 + off];
  t.h = p[6 + off];
  t.r1 = p[7 + off];
  t.r2 = p[8 + off];
  t.pi_factor = 3.14159265358979f / (t.r2 - t.r1);
  t.minus_half_over_n = (float)(-0.5 / (double)t.en);
  return t;
}

// Rigid ion, src/ri.cu:26-74.
constexpr float RI_ALPHA = 0.2f;
constexpr float RI_ALPHA_SQ = 0.04f;
constexpr float RI_PI_FACTOR = 0.225675833419103f; // 2 alpha / sqrt(pi)
constexpr float RI_K_C = 14.399645f;               // 1 / (4 pi eps0) in eV A

struct RIParams {
  float a[3], b[3], c[3], qq[3];
  float v_rc, dv_rc, cutoff;
  float mR, kR, mA, kA, r0; // generalised Morse on unlike pairs
};

__host__ __device__ inline RIParams ri_params(const float* p)
{
  RIParams q;
  const float q0 = p[0];
  q.qq[0] = q0 * q0 * RI_K_C;
  q.qq[1] = -0.5f * q0 * q0 * RI_K_C;
  q.qq[2] = 0.25f * q0 * q0 * RI_K_C;
  for (int t = 0; t < 3; ++t) {
    q.a[t] = p[1 + 3 * t];
    q.b[t] = 1.0f / p[2 + 3 * t];
    q.c[t] = p[3 + 3 * t];
  }
  q.cutoff = p[10];
  const float d0 = p[11], alpha = p[12], s = p[14];
  q.r0 = p[13];
  q.mR = d0 / (s - 1.0f);
  q.kR = -sqrtf(2.0f * s) * alpha;
  q.mA = s * d0 / (s - 1.0f);
  q.kA = -sqrtf(2.0f / s) * alpha;
  const float rc = q.cutoff;
  const float e = erfcf(RI_ALPHA * rc);
  q.v_rc = e / rc;
  q.dv_rc = -e / (rc * rc) - RI_PI_FACTOR * expf(-RI_ALPHA_SQ * rc * rc) / rc;
  return q;
}

// NN2B fixed cutoff taper, src/nn2b.cuh:36-38.
constexpr float NN2B_R1 = 7.0f;
constexpr float NN2B_R2 = 8.0f;
constexpr float NN2B_PI_FACTOR = 3.1415927f;

} // namespace gpuga
