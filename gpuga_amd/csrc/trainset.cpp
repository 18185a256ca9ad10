// Readers, neighbour list and geometry tables (host side of the set-up layer).
#include "trainset.h"

#include "common.h"

#include <algorithm>
#include <limits>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace gpuga {

static void set_box(float* h, const float* t);

namespace {

// Cursor over a whole file with fscanf-like %d / %f / %s conversions: skip
// white space, convert the longest valid prefix, fail when nothing converts.
class Scanner {
public:
  explicit Scanner(const std::string& path)
  {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) fail(strfmt("Failed to open %s!", path.c_str()));
    const std::streamoff n = f.tellg();
    text_.resize(n > 0 ? (size_t)n : 0);
    f.seekg(0);
    if (n > 0 && !f.read(&text_[0], n)) fail(strfmt("Failed to read %s!", path.c_str()));
    p_ = text_.c_str();
  }
  bool get_int(int& v)
  {
    skip();
    char* end;
    errno = 0;
    long x = std::strtol(p_, &end, 10);
    if (end == p_) return false;
    p_ = end;
    v = (int)x;
    return true;
  }
  bool get_float(float& v)
  {
    skip();
    char* end;
    float x = std::strtof(p_, &end);
    if (end == p_) return false;
    p_ = end;
    v = x;
    return true;
  }
  bool get_word(std::string& w)
  {
    skip();
    const char* b = p_;
    while (*p_ && !std::isspace((unsigned char)*p_)) ++p_;
    if (p_ == b) return false;
    w.assign(b, p_);
    return true;
  }

  const std::string& text() const { return text_; }
  size_t offset() const { return (size_t)(p_ - text_.c_str()); }

private:
  void skip()
  {
    while (*p_ && std::isspace((unsigned char)*p_)) ++p_;
  }
  std::string text_;
  const char* p_;
};

inline bool is_space(char c) { return std::isspace((unsigned char)c) != 0; }

// The body of train.in (after Nc, Nc_force and the Na line) read by 16
// threads: the text is cut at white space into pieces, each piece counts its
// tokens, and with the prefix sums each piece knows which field of which
// configuration its first token is, then converts its tokens in place.
// Returns false when the text does not hold exactly whole tokens of the
// expected kinds (a token that does not convert entirely, too few tokens):
// the caller then reads sequentially, which reports the reference's error.
bool read_body_parallel(const std::string& text, size_t begin, TrainSet& ts)
{
  const int Nc = ts.Nc;
  const size_t N = ts.N;
  // tokens per configuration: energy and virial (energy configurations),
  // box, then type x y z (+ fx fy fz for force configurations) per atom
  std::vector<long long> t0(Nc + 1, 0);
  for (int c = 0; c < Nc; ++c)
    t0[c + 1] = t0[c] + (c >= ts.Nc_force ? 7 : 0) + 9 + (long long)ts.Na[c] * (c < ts.Nc_force ? 7 : 4);
  const long long need = t0[Nc];
  const char* s = text.c_str();
  const size_t end = text.size();
  // pieces of at least 64 KiB, 16 per thread (parallel_for's grain)
  const int pieces = (int)std::max<size_t>(1, std::min<size_t>(256, (end - begin) / (1 << 16)));
  std::vector<size_t> cut(pieces + 1, end);
  cut[0] = begin;
  for (int k = 1; k < pieces; ++k) {
    size_t x = begin + (end - begin) * k / pieces;
    while (x < end && !is_space(s[x])) ++x; // a piece starts at white space
    cut[k] = std::max(x, cut[k - 1]);
  }
  std::vector<long long> count(pieces, 0);
  parallel_for(pieces, [&](int k) {
    long long n = 0;
    bool in = false;
    for (size_t x = cut[k]; x < cut[k + 1]; ++x) {
      const bool sp = is_space(s[x]);
      n += !sp && !in;
      in = !sp;
    }
    count[k] = n;
  });
  std::vector<long long> first(pieces + 1, 0);
  for (int k = 0; k < pieces; ++k) first[k + 1] = first[k] + count[k];
  if (first[pieces] < need) return false;
  std::vector<char> ok(pieces, 1);
  std::vector<float> box((size_t)Nc * 9);
  parallel_for(pieces, [&](int k) {
    long long g = first[k];
    if (g >= need) return;
    int c = (int)(std::upper_bound(t0.begin(), t0.end(), g) - t0.begin()) - 1;
    size_t x = cut[k];
    while (g < need && g < first[k + 1]) {
      while (x < cut[k + 1] && is_space(s[x])) ++x;
      size_t e = x;
      while (e < end && !is_space(s[e])) ++e;
      while (g >= t0[c + 1]) ++c;
      long long f = g - t0[c];
      const bool energy = c >= ts.Nc_force;
      char* stop = nullptr;
      if (energy && f < 7) {
        const float v = std::strtof(s + x, &stop);
        if (f == 0) ts.pe_ref[c] = v;
        else ts.virial_ref[c + (size_t)Nc * (f - 1)] = v;
      } else {
        f -= energy ? 7 : 0;
        if (f < 9) {
          box[(size_t)c * 9 + f] = std::strtof(s + x, &stop);
        } else {
          f -= 9;
          const int per = energy ? 4 : 7;
          const size_t i = ts.Na_sum[c] + f / per;
          const int q = (int)(f % per);
          if (q == 0) {
            errno = 0;
            ts.type[i] = (int)std::strtol(s + x, &stop, 10);
          } else if (q < 4) {
            ts.r[i + N * (q - 1)] = std::strtof(s + x, &stop);
          } else {
            ts.force_ref[i + N * (q - 4)] = std::strtof(s + x, &stop);
          }
        }
      }
      if (stop != s + e) {
        ok[k] = 0;
        return;
      }
      x = e;
      ++g;
    }
  });
  for (char o : ok)
    if (!o) return false;
  for (int c = 0; c < Nc; ++c) set_box(&ts.h[18 * (size_t)c], &box[(size_t)c * 9]);
  return true;
}

} // namespace

PotentialSpec read_potential(const std::string& dir)
{
  // src/fitness.cu:220-314
  PotentialSpec s;
  Scanner in(dir + "/potential.in");
  std::string name;
  const char* bad = "reading error for potential.in.";
  if (!in.get_word(name) || !in.get_int(s.type)) fail(bad);
  switch (s.type) {
  case POT_RI: s.V = 15; break;
  case POT_TERSOFF1: s.V = 9; break;
  case POT_TERSOFF2: s.V = 13; break;
  case POT_NN2B:
  case POT_NN2B_SPECIES:
    if (!in.get_word(name) || !in.get_int(s.num_neurons)) fail(bad);
    if (s.num_neurons < 1) fail("num_neurons should >= 1.");
    s.V = s.num_neurons * (s.num_neurons + 4) + 1 +
          (s.type == POT_NN2B_SPECIES ? s.num_neurons : 0);
    break;
  default: fail("unsupported potential type.");
  }
  if (!in.get_word(name) || !in.get_float(s.cutoff)) fail(bad);
  if (!in.get_word(name) || !in.get_float(s.w_force)) fail(bad);
  if (s.w_force < 0) fail("weight for force should >= 0");
  if (!in.get_word(name) || !in.get_float(s.w_energy)) fail(bad);
  if (s.w_energy < 0) fail("weight for energy should >= 0");
  if (!in.get_word(name) || !in.get_float(s.w_stress)) fail(bad);
  if (s.w_stress < 0) fail("weight for stress should >= 0");
  s.pmin.resize(s.V);
  s.pmax.resize(s.V);
  s.names.resize(s.V);
  for (int n = 0; n < s.V; ++n) {
    if (s.type < POT_NN2B) {
      if (!in.get_word(s.names[n]) || !in.get_float(s.pmin[n]) || !in.get_float(s.pmax[n]))
        fail(bad);
    } else {
      s.pmin[n] = -3.0f; // src/fitness.cu:308-309
      s.pmax[n] = 3.0f;
      s.names[n] = strfmt("w%d", n);
    }
  }
  return s;
}

#pragma clang fp contract(off)
void box_inverse(float* h)
{
  // src/fitness.cu:40-57, every product and difference rounded to float
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

static void set_box(float* h, const float* t)
{
  // file rows a, b, c become the columns of h (src/fitness.cu:115-123)
  h[0] = t[0]; h[3] = t[1]; h[6] = t[2];
  h[1] = t[3]; h[4] = t[4]; h[7] = t[5];
  h[2] = t[6]; h[5] = t[7]; h[8] = t[8];
  box_inverse(h);
}
#pragma clang fp contract(on)

static void check_counts(int Nc, int Nc_force)
{
  // src/fitness.cu:169-177
  if (Nc < 2) fail("Number of configurations should >= 2");
  if (Nc_force < 1) fail("Number of force configurations should >= 1");
  if (Nc_force > Nc - 1) fail("Number of potential configurations should >= 1");
}

static void finish_counts(TrainSet& ts)
{
  // src/fitness.cu:184-218
  ts.Na_sum.assign(ts.Nc, 0);
  ts.N = 0;
  ts.max_Na = 0;
  for (int c = 0; c < ts.Nc; ++c) {
    if (ts.Na[c] < 2) fail(strfmt("Number of atoms %d should >= 2", ts.Na[c]));
    ts.Na_sum[c] = ts.N;
    ts.N += ts.Na[c];
    ts.max_Na = std::max(ts.max_Na, ts.Na[c]);
  }
  ts.N_force = 0;
  for (int c = 0; c < ts.Nc_force; ++c) ts.N_force += ts.Na[c];
}

void finish_normalisers(TrainSet& ts)
{
  // src/fitness.cu:86-89,103-104,138-140,154-160, accumulated in FP64
  double emin = 0.0;
  for (int c = ts.Nc_force; c < ts.Nc; ++c) emin = std::min(emin, (double)ts.pe_ref[c]);
  ts.energy_minimum = emin;
  double fs = 0, ps = 0, vs = 0;
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < ts.N_force; ++i) {
      double v = ts.force_ref[i + (size_t)ts.N * k];
      fs += v * v;
    }
  for (int c = ts.Nc_force; c < ts.Nc; ++c) {
    double e = (double)ts.pe_ref[c] - emin;
    ps += e * e;
    for (int k = 0; k < 6; ++k) {
      double v = ts.virial_ref[c + (size_t)ts.Nc * k];
      vs += v * v;
    }
  }
  ts.force_square_sum = fs;
  ts.potential_square_sum = ps;
  ts.virial_square_sum = vs;
}

TrainSet read_train(const std::string& dir)
{
  // src/fitness.cu:59-161
  TrainSet ts;
  Scanner in(dir + "/train.in");
  const char* bad = "reading error for train.in.";
  if (!in.get_int(ts.Nc) || !in.get_int(ts.Nc_force)) fail("Reading error for train.in.");
  check_counts(ts.Nc, ts.Nc_force);
  ts.Na.resize(ts.Nc);
  for (int c = 0; c < ts.Nc; ++c)
    if (!in.get_int(ts.Na[c])) fail("Reading error for train.in.");
  finish_counts(ts);
  const int N = ts.N, Nc = ts.Nc;
  ts.h.assign((size_t)Nc * 18, 0.0f);
  ts.pe_ref.assign(Nc, 0.0f);
  ts.virial_ref.assign((size_t)Nc * 6, 0.0f);
  ts.type.assign(N, 0);
  ts.r.assign((size_t)N * 3, 0.0f);
  ts.force_ref.assign((size_t)N * 3, 0.0f);
  if (read_body_parallel(in.text(), in.offset(), ts)) {
    finish_normalisers(ts);
    return ts;
  }
  for (int c = 0; c < Nc; ++c) {
    if (c >= ts.Nc_force) {
      if (!in.get_float(ts.pe_ref[c])) fail(bad);
      for (int k = 0; k < 6; ++k)
        if (!in.get_float(ts.virial_ref[c + (size_t)Nc * k])) fail(bad);
    }
    float t[9];
    for (int k = 0; k < 9; ++k)
      if (!in.get_float(t[k])) fail(bad);
    set_box(&ts.h[18 * (size_t)c], t);
    for (int k = 0; k < ts.Na[c]; ++k) {
      const size_t i = ts.Na_sum[c] + k;
      if (!in.get_int(ts.type[i]) || !in.get_float(ts.r[i]) || !in.get_float(ts.r[i + N]) ||
          !in.get_float(ts.r[i + 2 * (size_t)N]))
        fail(bad);
      if (c < ts.Nc_force) {
        if (!in.get_float(ts.force_ref[i]) || !in.get_float(ts.force_ref[i + N]) ||
            !in.get_float(ts.force_ref[i + 2 * (size_t)N]))
          fail(bad);
      }
    }
  }
  finish_normalisers(ts);
  return ts;
}

TrainSet make_train(int Nc, int Nc_force, const int* Na, const int* type, const float* box9,
                    const float* r, const float* pe_ref, const float* virial_ref,
                    const float* force_ref)
{
  TrainSet ts;
  check_counts(Nc, Nc_force);
  ts.Nc = Nc;
  ts.Nc_force = Nc_force;
  ts.Na.assign(Na, Na + Nc);
  finish_counts(ts);
  const size_t N = ts.N;
  ts.type.assign(type, type + N);
  ts.r.assign(r, r + 3 * N);
  ts.force_ref.assign(3 * N, 0.0f);
  if (force_ref) std::copy(force_ref, force_ref + 3 * N, ts.force_ref.begin());
  ts.pe_ref.assign(Nc, 0.0f);
  if (pe_ref) std::copy(pe_ref, pe_ref + Nc, ts.pe_ref.begin());
  ts.virial_ref.assign((size_t)Nc * 6, 0.0f);
  if (virial_ref) std::copy(virial_ref, virial_ref + 6 * (size_t)Nc, ts.virial_ref.begin());
  ts.h.assign((size_t)Nc * 18, 0.0f);
  for (int c = 0; c < Nc; ++c) set_box(&ts.h[18 * (size_t)c], box9 + 9 * (size_t)c);
  finish_normalisers(ts);
  return ts;
}

int config_of(const TrainSet& ts, int i)
{
  auto it = std::upper_bound(ts.Na_sum.begin(), ts.Na_sum.end(), i);
  return (int)(it - ts.Na_sum.begin()) - 1;
}

#pragma clang fp contract(off)
// FP32 minimum image exactly as src/mic.cuh:20-32 is written.
static inline float mic_d2(const float* h, float x, float y, float z)
{
  float sx = h[9] * x + h[10] * y + h[11] * z;
  float sy = h[12] * x + h[13] * y + h[14] * z;
  float sz = h[15] * x + h[16] * y + h[17] * z;
  sx -= std::nearbyint(sx);
  sy -= std::nearbyint(sy);
  sz -= std::nearbyint(sz);
  float rx = h[0] * sx + h[1] * sy + h[2] * sz;
  float ry = h[3] * sx + h[4] * sy + h[5] * sz;
  float rz = h[6] * sx + h[7] * sy + h[8] * sz;
  return rx * rx + ry * ry + rz * rz;
}

NeighborList build_neighbors(const TrainSet& ts, float cutoff)
{
  // src/neighbor.cu:30-66 (and rc^2 in float, :72); CSR instead of NL[k*N+i]
  NeighborList nl;
  const int N = ts.N;
  const float rc2 = cutoff * cutoff;
  nl.NN.assign(N, 0);
  nl.start.assign(N, 0);
  std::vector<std::vector<int>> tmp(N);
  parallel_for(ts.Nc, [&](int c) {
    const float* h = &ts.h[18 * (size_t)c];
    const int b = ts.Na_sum[c], e = b + ts.Na[c];
    for (int i = b; i < e; ++i) {
      const float xi = ts.r[i], yi = ts.r[i + N], zi = ts.r[i + 2 * (size_t)N];
      for (int j = b; j < e; ++j) {
        if (j == i) continue;
        float x = ts.r[j] - xi, y = ts.r[j + N] - yi, z = ts.r[j + 2 * (size_t)N] - zi;
        if (mic_d2(h, x, y, z) < rc2) tmp[i].push_back(j);
      }
    }
  });
  long long total = 0;
  for (int i = 0; i < N; ++i) {
    nl.NN[i] = (int)tmp[i].size();
    nl.start[i] = (int)total;
    total += nl.NN[i];
    nl.max_nn = std::max(nl.max_nn, nl.NN[i]);
  }
  if (total > 0x7fffffffLL) fail("neighbour list too large");
  nl.total = (int)total;
  nl.NL.resize(total);
  for (int i = 0; i < N; ++i) std::copy(tmp[i].begin(), tmp[i].end(), nl.NL.begin() + nl.start[i]);
  nl.rev.resize(total);
  parallel_for(N, [&](int i) {
    for (int s = 0; s < nl.NN[i]; ++s) {
      const int j = nl.NL[nl.start[i] + s];
      int off = 0;
      for (int u = 0; u < nl.NN[j]; ++u)
        if (nl.NL[nl.start[j] + u] == i) {
          off = u;
          break;
        }
      nl.rev[nl.start[i] + s] = nl.start[j] + off;
    }
  });
  return nl;
}
#pragma clang fp contract(on)

Geometry build_geometry(const TrainSet& ts, const NeighborList& nl, bool triplets)
{
  Geometry g;
  const int N = ts.N;
  g.bond.resize(nl.total);
  std::vector<double> rx(nl.total), ry(nl.total), rz(nl.total), rd(nl.total);
  parallel_for(ts.Nc, [&](int c) {
    const float* h = &ts.h[18 * (size_t)c];
    const int b = ts.Na_sum[c], e = b + ts.Na[c];
    for (int i = b; i < e; ++i)
      for (int s = 0; s < nl.NN[i]; ++s) {
        const int k = nl.start[i] + s, j = nl.NL[k];
        // the minimum image of src/mic.cuh:20-32 evaluated in FP64 on the float box
        double x = (double)ts.r[j] - ts.r[i];
        double y = (double)ts.r[j + N] - ts.r[i + N];
        double z = (double)ts.r[j + 2 * (size_t)N] - ts.r[i + 2 * (size_t)N];
        double sx = (double)h[9] * x + (double)h[10] * y + (double)h[11] * z;
        double sy = (double)h[12] * x + (double)h[13] * y + (double)h[14] * z;
        double sz = (double)h[15] * x + (double)h[16] * y + (double)h[17] * z;
        sx -= std::nearbyint(sx);
        sy -= std::nearbyint(sy);
        sz -= std::nearbyint(sz);
        x = h[0] * sx + h[1] * sy + h[2] * sz;
        y = h[3] * sx + h[4] * sy + h[5] * sz;
        z = h[6] * sx + h[7] * sy + h[8] * sz;
        const double d = std::sqrt(x * x + y * y + z * z);
        rx[k] = x; ry[k] = y; rz[k] = z; rd[k] = d;
        g.bond[k] = Vec4{(float)x, (float)y, (float)z, (float)d};
      }
  });
  if (!triplets) return g;
  long long total = 0;
  g.tri_start = triplet_offsets(nl, &total);
  g.tri_total = total;
  g.tri.assign(total, Vec4{0, 0, 0, 0});
  parallel_for(N, [&](int i) {
    const int nn = nl.NN[i], b0 = nl.start[i];
    for (int s = 0; s < nn; ++s)
      for (int t = 0; t < nn; ++t) {
        if (s == t) continue;
        const int ks = b0 + s, kt = b0 + t;
        // cos(theta) between r_ij and r_ik and its gradient with respect to r_ij
        // (the factors of src/minimal_tersoff.cu:291-306)
        const double inv = 1.0 / (rd[ks] * rd[kt]);
        const double cs = (rx[ks] * rx[kt] + ry[ks] * ry[kt] + rz[ks] * rz[kt]) * inv;
        const double cd = cs / (rd[ks] * rd[ks]);
        g.tri[g.tri_start[i] + s * nn + t] =
          Vec4{(float)cs, (float)(rx[kt] * inv - rx[ks] * cd), (float)(ry[kt] * inv - ry[ks] * cd),
               (float)(rz[kt] * inv - rz[ks] * cd)};
      }
  });
  return g;
}

std::vector<int> triplet_offsets(const NeighborList& nl, long long* total)
{
  std::vector<int> off(nl.NN.size());
  long long t = 0;
  for (size_t i = 0; i < nl.NN.size(); ++i) {
    off[i] = (int)t;
    t += (long long)nl.NN[i] * nl.NN[i];
    if (t > 0x7fffffffLL) fail("triplet table too large");
  }
  *total = t;
  return off;
}

std::vector<int> bond_types(const TrainSet& ts, const NeighborList& nl, bool three_body)
{
  std::vector<int> p(nl.total);
  for (int i = 0; i < ts.N; ++i)
    for (int s = 0; s < nl.NN[i]; ++s) {
      const int b = nl.start[i] + s, j = nl.NL[b];
      p[b] = three_body ? (ts.type[i] != ts.type[j] ? 1 : 0) : ts.type[i] + ts.type[j];
    }
  return p;
}

int record_slots(int max_nn)
{
  return max_nn <= 4 ? 4 : max_nn <= 8 ? 8 : max_nn <= 16 ? 16 : 0;
}

float tersoff_reach(const PotentialSpec& s)
{
  if ((s.type != POT_TERSOFF1 && s.type != POT_TERSOFF2) || s.V < 1)
    return std::numeric_limits<float>::infinity();
  // the kernels' denormalisation lo + (hi - lo) * g, unfused, at g = 1
  volatile float lo = s.pmin[s.V - 1], hi = s.pmax[s.V - 1];
  volatile float b = hi - lo;
  volatile float bx = b * 1.0f;
  return lo + bx;
}

NeighborList prune_neighbors(const NeighborList& nl, const std::vector<float>& d, float reach)
{
  const int N = (int)nl.NN.size();
  NeighborList p;
  p.NN.assign(N, 0);
  p.start.assign(N, 0);
  std::vector<int> map(nl.total, -1); // full bond index -> pruned index
  long long total = 0;
  for (int i = 0; i < N; ++i) {
    p.start[i] = (int)total;
    for (int b = nl.start[i]; b < nl.start[i] + nl.NN[i]; ++b)
      if (d[b] < reach) map[b] = (int)total++;
    p.NN[i] = (int)total - p.start[i];
    p.max_nn = std::max(p.max_nn, p.NN[i]);
  }
  p.total = (int)total;
  p.NL.resize(total);
  p.rev.resize(total);
  for (int b = 0; b < nl.total; ++b) {
    if (map[b] < 0) continue;
    const int r = map[nl.rev[b]];
    if (r < 0) fail("internal: pruned neighbour list is not symmetric");
    p.NL[map[b]] = nl.NL[b];
    p.rev[map[b]] = r;
  }
  return p;
}

std::vector<Vec4> build_records(const NeighborList& nl, const Geometry& g,
                                const std::vector<int>& ptype, int M)
{
  const int N = (int)nl.NN.size();
  const size_t RS = 1 + M + (size_t)M * M + M / 2;
  std::vector<Vec4> rec((size_t)N * RS, Vec4{0, 0, 0, 0});
  parallel_for(N, [&](int i) {
    const int nn = nl.NN[i], b0 = nl.start[i];
    Vec4* R = &rec[(size_t)i * RS];
    int hdr[4] = {nn, 0, b0, 0};
    float dmax = 0.0f; // the longest bond, in the header's last word
    for (int u = 0; u < nn; ++u) {
      hdr[1] |= (ptype[b0 + u] & 1) << u;
      R[1 + u] = g.bond[b0 + u];
      dmax = std::max(dmax, g.bond[b0 + u].w);
      for (int w = 0; w < nn; ++w)
        if (w != u) R[1 + M + u * M + w] = g.tri[g.tri_start[i] + u * nn + w];
    }
    std::memcpy(&hdr[3], &dmax, sizeof dmax);
    std::memcpy(&R[0], hdr, sizeof hdr);
    int* rv = reinterpret_cast<int*>(&R[1 + M + M * M]);
    for (int u = 0; u < nn; ++u) rv[u] = nl.rev[b0 + u];
    float* inv = reinterpret_cast<float*>(&R[1 + M + M * M + M / 4]);
    for (int u = 0; u < nn; ++u) inv[u] = (float)(1.0 / (double)g.bond[b0 + u].w);
  });
  return rec;
}

} // namespace gpuga
