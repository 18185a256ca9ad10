// Fitness of one parameter set from per-atom force, virial and energy:
// per-configuration sums in FP64 and the three relative RMS errors.
// Restates the reductions of src/fitness.cu:392-555 (gpu_sum_force_error,
// gpu_sum_pe_error and the host loops) as two deterministic launches.
#include "kernels.h"

namespace gpuga {

namespace {

constexpr int RB = 256;

// Tree sum of one double per thread over a 256-thread block, fixed order.
__device__ double block_sum(double v, double* sh)
{
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if (t < o) sh[t] += sh[t + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(RB) config_terms(DevSet s, const float* __restrict__ f,
                                                   const float* __restrict__ virial,
                                                   const float* __restrict__ pe,
                                                   double* __restrict__ cerr)
{
  __shared__ double sh[RB];
  const int c = blockIdx.x;
  const int a0 = s.Na_sum[c], na = s.Na[c];
  const size_t N = s.N;
  if (c < s.Nc_force) {
    double e = 0.0;
    for (int k = threadIdx.x; k < na; k += RB) {
      const int i = a0 + k;
      const float dx = f[i] - s.force_ref[i];
      const float dy = f[i + N] - s.force_ref[i + N];
      const float dz = f[i + 2 * N] - s.force_ref[i + 2 * N];
      e += (double)dx * dx + (double)dy * dy + (double)dz * dz;
    }
    e = block_sum(e, sh);
    if (threadIdx.x == 0) {
      cerr[3 * c + 0] = 0.0;
      cerr[3 * c + 1] = 0.0;
      cerr[3 * c + 2] = e;
    }
  } else {
    double q[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < na; k += RB) {
      const int i = a0 + k;
      q[0] += pe[i];
#pragma unroll
      for (int u = 0; u < 6; ++u) q[1 + u] += virial[i + N * u];
    }
    double sum[7];
#pragma unroll
    for (int u = 0; u < 7; ++u) sum[u] = block_sum(q[u], sh);
    if (threadIdx.x == 0) {
      const double dE = sum[0] - (double)s.pe_ref[c];
      double eW = 0.0;
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        const double d = sum[1 + u] - (double)s.virial_ref[c + (size_t)s.Nc * u];
        eW += d * d;
      }
      cerr[3 * c + 0] = dE * dE;
      cerr[3 * c + 1] = eW;
      cerr[3 * c + 2] = 0.0;
    }
  }
}

__global__ void __launch_bounds__(RB) finish_terms(int Nc, FitnessNorm nrm,
                                                   const double* __restrict__ cerr,
                                                   float* __restrict__ fitness,
                                                   double* __restrict__ terms)
{
  __shared__ double sh[RB];
  double q[3] = {0, 0, 0};
  for (int c = threadIdx.x; c < Nc; c += RB) {
    q[0] += cerr[3 * c + 0];
    q[1] += cerr[3 * c + 1];
    q[2] += cerr[3 * c + 2];
  }
  const double se = block_sum(q[0], sh);
  const double sv = block_sum(q[1], sh);
  const double sf = block_sum(q[2], sh);
  if (threadIdx.x == 0) {
    const double tE = sqrt(se * nrm.inv_pss), tW = sqrt(sv * nrm.inv_vss),
                 tF = sqrt(sf * nrm.inv_fss);
    terms[0] = tE;
    terms[1] = tW;
    terms[2] = tF;
    fitness[0] = (float)(nrm.w_energy * tE + nrm.w_stress * tW + nrm.w_force * tF);
  }
}

} // namespace

void launch_fitness_terms(const DevSet& s, const FitnessNorm& nrm, const float* f,
                          const float* virial, const float* pe, double* cerr, float* fitness,
                          double* terms, hipStream_t st)
{
  config_terms<<<s.Nc, RB, 0, st>>>(s, f, virial, pe, cerr);
  finish_terms<<<1, RB, 0, st>>>(s.Nc, nrm, cerr, fitness, terms);
  GPUGA_HIP(hipGetLastError());
}

} // namespace gpuga
