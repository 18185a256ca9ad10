// Neighbour list on the device (the reference builds it on the GPU as well,
// src/neighbor.cu:30-66), in the CSR form of build_neighbors (trainset.cpp)
// and bit-identical to it: the minimum-image distance is the same FP32
// expression evaluated in the same order with no fused multiply-add (as
// src/mic.cuh:20-32 is written), rounding half to even.
//
// Three launches, one thread per atom: count the neighbours (d^2 < rc^2, j in
// the atom's configuration, ascending), write them at the offsets of the
// exclusive prefix sum of the counts (computed on the host, which needs the
// list anyway), then find each bond's reverse bond.
//
// The parameter-independent geometry tables follow on the device as well,
// bit-identical to build_geometry / build_records (trainset.cpp): the same
// FP64 expressions without contraction (IEEE double add, multiply, divide and
// square root are correctly rounded on both sides), rounded to float.
#include <hip/hip_runtime.h>

#include "common.h"
#include "trainset.h"

namespace gpuga {

namespace {

#pragma clang fp contract(off)

struct NLView {
  int N;
  float rc2;
  const float* r;      // [3][N]
  const float* h;      // [Nc][18] box columns, then the inverse
  const int* cfg;      // [N] configuration of atom i
  const int* Na_sum;   // [Nc]
  const int* Na;       // [Nc]
};

__device__ __forceinline__ float mic_d2(const float* h, float x, float y, float z)
{
  float sx = h[9] * x + h[10] * y + h[11] * z;
  float sy = h[12] * x + h[13] * y + h[14] * z;
  float sz = h[15] * x + h[16] * y + h[17] * z;
  sx -= __builtin_rintf(sx);
  sy -= __builtin_rintf(sy);
  sz -= __builtin_rintf(sz);
  const float rx = h[0] * sx + h[1] * sy + h[2] * sz;
  const float ry = h[3] * sx + h[4] * sy + h[5] * sz;
  const float rz = h[6] * sx + h[7] * sy + h[8] * sz;
  return rx * rx + ry * ry + rz * rz;
}

// FILL = false: NN[i] = number of neighbours; true: NL[start[i] + k] = the
// k-th neighbour
template <bool FILL>
__global__ void __launch_bounds__(256) nl_pass(NLView v, int* NN, const int* start, int* NL)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.N) return;
  const int c = v.cfg[i];
  // the box in registers (read through the pointer, it is loaded again for
  // every pair)
  float h[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) h[k] = v.h[18 * (long)c + k];
  const int b = v.Na_sum[c], e = b + v.Na[c];
  const long N = v.N;
  const float xi = v.r[i], yi = v.r[i + N], zi = v.r[i + 2 * N];
  int k = 0;
  int* out = FILL ? NL + start[i] : nullptr;
  for (int j = b; j < e; ++j) {
    if (j == i) continue;
    const float x = v.r[j] - xi, y = v.r[j + N] - yi, z = v.r[j + 2 * N] - zi;
    if (mic_d2(h, x, y, z) < v.rc2) {
      if (FILL) out[k] = j;
      ++k;
    }
  }
  if (!FILL) NN[i] = k;
}

// rev[b] for bond b = (i, j): the slot of i in j's list, or j's first slot
// when i is absent (lists are ascending and without repeats, so a binary
// search finds the slot the linear search of the host build finds)
__global__ void __launch_bounds__(256) nl_reverse(int N, const int* NN, const int* start,
                                                  const int* NL, int* rev)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int s0 = start[i], n = NN[i];
  for (int s = 0; s < n; ++s) {
    const int j = NL[s0 + s];
    const int t0 = start[j];
    int lo = 0, hi = NN[j]; // first slot with NL >= i
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (NL[t0 + mid] < i) lo = mid + 1;
      else hi = mid;
    }
    const int off = lo < NN[j] && NL[t0 + lo] == i ? lo : 0;
    rev[s0 + s] = t0 + off;
  }
}

// Minimum image of bond k = (i, j) in FP64 on the float box, as
// build_geometry: rb[k] = (x, y, z, |r|), bond[k] = the same rounded to float.
__global__ void __launch_bounds__(256) geo_bonds(NLView v, const int* NN, const int* start,
                                                 const int* NL, double4* rb, float4* bond)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.N) return;
  const float* h = v.h + 18 * (long)v.cfg[i];
  const long N = v.N;
  const int n = NN[i], b0 = start[i];
  for (int s = 0; s < n; ++s) {
    const int k = b0 + s, j = NL[k];
    double x = (double)v.r[j] - (double)v.r[i];
    double y = (double)v.r[j + N] - (double)v.r[i + N];
    double z = (double)v.r[j + 2 * N] - (double)v.r[i + 2 * N];
    double sx = (double)h[9] * x + (double)h[10] * y + (double)h[11] * z;
    double sy = (double)h[12] * x + (double)h[13] * y + (double)h[14] * z;
    double sz = (double)h[15] * x + (double)h[16] * y + (double)h[17] * z;
    sx -= __builtin_rint(sx);
    sy -= __builtin_rint(sy);
    sz -= __builtin_rint(sz);
    x = (double)h[0] * sx + (double)h[1] * sy + (double)h[2] * sz;
    y = (double)h[3] * sx + (double)h[4] * sy + (double)h[5] * sz;
    z = (double)h[6] * sx + (double)h[7] * sy + (double)h[8] * sz;
    const double d = __builtin_sqrt(x * x + y * y + z * z);
    rb[k] = make_double4(x, y, z, d);
    bond[k] = make_float4((float)x, (float)y, (float)z, (float)d);
  }
}

// cos(theta_jik) and its gradient with respect to r_ij for every ordered
// pair (s, t) of atom i's bonds, zero where s == t
__global__ void __launch_bounds__(256) geo_triplets(int N, const int* NN, const int* start,
                                                    const int* tri0, const double4* rb, float4* tri)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int nn = NN[i], b0 = start[i];
  float4* out = tri + tri0[i];
  for (int s = 0; s < nn; ++s) {
    const double4 a = rb[b0 + s];
    for (int t = 0; t < nn; ++t) {
      if (s == t) {
        out[s * nn + t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        continue;
      }
      const double4 b = rb[b0 + t];
      const double inv = 1.0 / (a.w * b.w);
      const double cs = (a.x * b.x + a.y * b.y + a.z * b.z) * inv;
      const double cd = cs / (a.w * a.w);
      out[s * nn + t] = make_float4((float)cs, (float)(b.x * inv - a.x * cd),
                                    (float)(b.y * inv - a.y * cd), (float)(b.z * inv - a.z * cd));
    }
  }
}

// One float4 (slot k) of atom i's record.
__device__ inline void geo_record_entry(int i, int k, int M, const int* NN, const int* start,
                                        const int* rev, const int* ptype, const int* tri0,
                                        const float4* bond, const float4* tri, float4* out)
{
  const int nn = NN[i], b0 = start[i];
  const int REV = 1 + M + M * M, INV = REV + M / 4;
  float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (k == 0) {
    int mask = 0;
    float dmax = 0.0f;
    for (int u = 0; u < nn; ++u) {
      mask |= (ptype[b0 + u] & 1) << u;
      dmax = fmaxf(dmax, bond[b0 + u].w);
    }
    v = make_float4(__int_as_float(nn), __int_as_float(mask), __int_as_float(b0), dmax);
  } else if (k <= M) {
    if (k - 1 < nn) v = bond[b0 + k - 1];
  } else if (k < REV) {
    const int u = (k - 1 - M) / M, w = (k - 1 - M) % M;
    if (u < nn && w < nn && u != w) v = tri[tri0[i] + u * nn + w];
  } else if (k < INV) {
    int r[4];
    for (int c = 0; c < 4; ++c) {
      const int u = 4 * (k - REV) + c;
      r[c] = u < nn ? rev[b0 + u] : 0;
    }
    v = make_float4(__int_as_float(r[0]), __int_as_float(r[1]), __int_as_float(r[2]),
                    __int_as_float(r[3]));
  } else {
    float f[4];
    for (int c = 0; c < 4; ++c) {
      const int u = 4 * (k - INV) + c;
      f[c] = u < nn ? (float)(1.0 / (double)bond[b0 + u].w) : 0.0f;
    }
    v = make_float4(f[0], f[1], f[2], f[3]);
  }
  *out = v;
}

// Atom records of the Tersoff population kernel, the layout of
// build_records: header (nn, species-pair mask, first bond, longest bond),
// M bonds, M x M triplets, M reverse bonds, M inverse distances,
// zero-padded.  One thread per float4 of the record array (grid-stride), so
// that a wave's stores are contiguous.
__global__ void __launch_bounds__(256) geo_records(int N, int M, const int* NN, const int* start,
                                                   const int* rev, const int* ptype, const int* tri0,
                                                   const float4* bond, const float4* tri, float4* rec)
{
  const int RS = 1 + M + M * M + M / 2;
  const long total = (long)N * RS;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / RS), k = (int)(e - (long)i * RS);
    geo_record_entry(i, k, M, NN, start, rev, ptype, tri0, bond, tri, rec + e);
  }
}

#pragma clang fp contract(on)

// Device copies of the positions, boxes and configuration index.
struct NLInput {
  DeviceBuffer<float> r, h;
  DeviceBuffer<int> cfg, Na_sum, Na, NN;
  NLView view{};
  NLInput(const TrainSet& ts, float cutoff, hipStream_t st)
  {
    std::vector<int> c(ts.N);
    for (int k = 0; k < ts.Nc; ++k)
      for (int i = ts.Na_sum[k]; i < ts.Na_sum[k] + ts.Na[k]; ++i) c[i] = k;
    r.upload(ts.r, st);
    h.upload(ts.h, st);
    cfg.upload(c, st);
    Na_sum.upload(ts.Na_sum, st);
    Na.upload(ts.Na, st);
    NN.resize(std::max(ts.N, 1));
    view.N = ts.N;
    view.rc2 = cutoff * cutoff;
    view.r = r.get();
    view.h = h.get();
    view.cfg = cfg.get();
    view.Na_sum = Na_sum.get();
    view.Na = Na.get();
    // the uploads read host memory that c owns
    GPUGA_HIP(hipStreamSynchronize(st));
  }
  int blocks() const { return (view.N + 255) / 256; }
  void count(hipStream_t st)
  {
    if (view.N == 0) return;
    nl_pass<false><<<blocks(), 256, 0, st>>>(view, NN.get(), nullptr, nullptr);
    GPUGA_HIP(hipGetLastError());
  }
};

} // namespace

std::vector<int> count_neighbors_device(const TrainSet& ts, float cutoff, hipStream_t st)
{
  NLInput in(ts, cutoff, st);
  in.count(st);
  std::vector<int> nn(ts.N);
  in.NN.download(nn.data(), nn.size(), st);
  GPUGA_HIP(hipStreamSynchronize(st));
  return nn;
}

void build_geometry_device(const TrainSet& ts, const DeviceNeighbors& d, float4* bond,
                           float4* tri, hipStream_t st)
{
  if (d.total == 0 || ts.N == 0) return;
  NLInput in(ts, 0.0f, st);
  DeviceBuffer<double4> rb(d.total);
  geo_bonds<<<in.blocks(), 256, 0, st>>>(in.view, d.NN, d.start, d.NL, rb.get(), bond);
  GPUGA_HIP(hipGetLastError());
  if (tri) {
    geo_triplets<<<in.blocks(), 256, 0, st>>>(ts.N, d.NN, d.start, d.tri0, rb.get(), tri);
    GPUGA_HIP(hipGetLastError());
  }
  // rb and the inputs are freed on return
  GPUGA_HIP(hipStreamSynchronize(st));
}

void build_records_device(int N, int M, const DeviceNeighbors& d, const int* ptype,
                          const float4* bond, const float4* tri, float4* rec, hipStream_t st)
{
  if (N == 0) return;
  if (M != 4 && M != 8 && M != 16) fail("atom records hold 4, 8 or 16 bonds");
  const long total = (long)N * (1 + M + (long)M * M + M / 2);
  // grid-stride: at most 2^20 workgroups, whatever the table size
  const long blocks = std::min<long>((total + 255) / 256, 1L << 20);
  geo_records<<<(unsigned)blocks, 256, 0, st>>>(N, M, d.NN, d.start, d.rev, ptype, d.tri0, bond,
                                                tri, rec);
  GPUGA_HIP(hipGetLastError());
}

NeighborList build_neighbors_device(const TrainSet& ts, float cutoff, hipStream_t st)
{
  NeighborList nl;
  const int N = ts.N;
  NLInput in(ts, cutoff, st);
  in.count(st);
  nl.NN.resize(N);
  in.NN.download(nl.NN.data(), N, st);
  GPUGA_HIP(hipStreamSynchronize(st));
  nl.start.resize(N);
  long long total = 0;
  for (int i = 0; i < N; ++i) {
    nl.start[i] = (int)total;
    total += nl.NN[i];
    nl.max_nn = std::max(nl.max_nn, nl.NN[i]);
  }
  if (total > 0x7fffffffLL) fail("neighbour list too large");
  nl.total = (int)total;
  nl.NL.resize(total);
  nl.rev.resize(total);
  if (total == 0) return nl;
  DeviceBuffer<int> start, NL, rev;
  start.upload(nl.start, st);
  NL.resize(total);
  rev.resize(total);
  nl_pass<true><<<in.blocks(), 256, 0, st>>>(in.view, nullptr, start.get(), NL.get());
  GPUGA_HIP(hipGetLastError());
  nl_reverse<<<in.blocks(), 256, 0, st>>>(N, in.NN.get(), start.get(), NL.get(), rev.get());
  GPUGA_HIP(hipGetLastError());
  NL.download(nl.NL.data(), total, st);
  rev.download(nl.rev.data(), total, st);
  GPUGA_HIP(hipStreamSynchronize(st));
  return nl;
}

} // namespace gpuga
