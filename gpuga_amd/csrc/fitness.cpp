// Fitness evaluator host code: set-up, plug-ins, population plans, predict.
#include "fitness.h"
#include "noalloc.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>

namespace gpuga {

int tersoff_population_occupancy(int cls, bool two, bool pack, int block);

// Neighbours within R2 the looped Tersoff population kernel takes (its LDS
// rows: 4 x 64 lanes x 4 B per neighbour and wave, 64 KiB at this bound).
constexpr int kLoopMaxNeighbors = 64;
int ri_population_occupancy(int block);
int nn2b_population_occupancy(int block);

// ------------------------------------------------------------ plug-ins

namespace {

// Host denormalisation with float arithmetic and no fused multiply-add,
// exactly as src/fitness.cu:319-324 (this file is compiled with
// -ffp-contract=off).
void denormalise_host(const PotentialSpec& s, const float* genes, float* out)
{
  for (int m = 0; m < s.V; ++m) {
    const float a = s.pmin[m];
    const float b = s.pmax[m] - a;
    const float bx = b * genes[m];
    out[m] = a + bx;
  }
}

class TersoffPotential : public Potential {
public:
  explicit TersoffPotential(int num_types) : num_types_(num_types) {}
  void initialize(int, int) override {}
  void init_bonds(int B)
  {
    b_.resize(std::max(B, 1));
    bp_.resize(std::max(B, 1));
    f12_.resize(3 * (size_t)std::max(B, 1));
  }
  void update_potential(const std::vector<float>& p) override
  {
    q_ = tersoff_params(num_types_, p.data());
  }
  void find_force(const DevSet& s, float* f, float* virial, float* pe, hipStream_t st,
                  float* ws) override
  {
    // workspace: b and b' per bond, then f_ij (3 per bond)
    const size_t B = b_.size();
    if (ws) launch_tersoff_plain(s, q_, ws, ws + B, ws + 2 * B, f, virial, pe, st);
    else launch_tersoff_plain(s, q_, b_.get(), bp_.get(), f12_.get(), f, virial, pe, st);
  }
  size_t workspace_floats() const override { return 5 * b_.size(); }

private:
  int num_types_;
  TersoffParams q_{};
  DeviceBuffer<float> b_, bp_, f12_;
};

class RIPotential : public Potential {
public:
  explicit RIPotential(const float4* tab) : tab_(tab) {}
  void initialize(int, int) override {}
  void update_potential(const std::vector<float>& p) override { q_ = ri_params(p.data()); }
  void find_force(const DevSet& s, float* f, float* virial, float* pe, hipStream_t st,
                  float*) override
  {
    launch_ri_plain(s, q_, tab_, f, virial, pe, st);
  }

private:
  const float4* tab_;
  RIParams q_{};
};

class NN2BPotential : public Potential {
public:
  // species: the first layer also takes the pair's species code
  NN2BPotential(int n, bool species) : n_(n), species_(species) {}
  void initialize(int, int) override { w_.resize(n_ * (n_ + 4) + 1 + (species_ ? n_ : 0)); }
  void update_potential(const std::vector<float>& p) override
  {
    host_ = p;
    dirty_ = true;
  }
  void find_force(const DevSet& s, float* f, float* virial, float* pe, hipStream_t st,
                  float*) override
  {
    if (dirty_) {
      w_.upload(host_, st);
      dirty_ = false;
    }
    launch_nn2b_plain(s, n_, species_, w_.get(), f, virial, pe, st);
  }

private:
  int n_;
  bool species_;
  std::vector<float> host_;
  bool dirty_ = false;
  DeviceBuffer<float> w_;
};

} // namespace

std::unique_ptr<Potential> make_potential(const PotentialSpec& spec, const DeviceSet& ds)
{
  switch (spec.type) {
  case POT_TERSOFF1:
  case POT_TERSOFF2: {
    auto p = std::make_unique<TersoffPotential>(spec.type == POT_TERSOFF2 ? 2 : 1);
    p->init_bonds((int)ds.bond.size());
    return p;
  }
  case POT_RI: return std::make_unique<RIPotential>(ds.ri_tab.get());
  case POT_NN2B:
  case POT_NN2B_SPECIES: {
    auto p = std::make_unique<NN2BPotential>(spec.num_neurons, spec.type == POT_NN2B_SPECIES);
    p->initialize(0, 0);
    return p;
  }
  }
  fail("unsupported potential type.");
}

size_t DeviceSet::bytes() const
{
  return Na.bytes() + Na_sum.bytes() + type.bytes() + nn.bytes() + bond0.bytes() + nl.bytes() +
         rev.bytes() + ptype.bytes() + tri0.bytes() + bond.bytes() + tri.bytes() +
         ri_tab.bytes() + rec.bytes() + pair_hdr.bytes() + pair_tab.bytes() + pe_ref.bytes() +
         virial_ref.bytes() + force_ref.bytes() + p_nn.bytes() + p_b0.bytes() + p_nl.bytes() +
         p_rev.bytes() + p_ptype.bytes() + p_tri0.bytes() + p_bond.bytes() + p_tri.bytes() +
         ri_cconst.bytes() + ri_fconst.bytes();
}

void PopPlan::drop_ga_graphs()
{
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < kGaGraphs; ++j) {
      if (ga_exec[k][j]) (void)hipGraphExecDestroy(ga_exec[k][j]);
      if (ga_graph[k][j]) (void)hipGraphDestroy(ga_graph[k][j]);
      ga_exec[k][j] = nullptr;
      ga_graph[k][j] = nullptr;
    }
}

void PopPlan::drop_share_graphs()
{
  for (int j = 0; j < kShares; ++j) {
    if (share_exec[j]) (void)hipGraphExecDestroy(share_exec[j]);
    if (share_graph[j]) (void)hipGraphDestroy(share_graph[j]);
    share_exec[j] = nullptr;
    share_graph[j] = nullptr;
  }
}

PopPlan::~PopPlan()
{
  drop_share_graphs();
  drop_ga_graphs();
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
}

// ------------------------------------------------------------ set-up

Fitness::Fitness(const std::string& dir, const FitnessOptions& opt) : opt_(opt)
{
  spec_ = read_potential(dir);
  ts_ = read_train(dir);
  setup();
}

Fitness::Fitness(const PotentialSpec& spec, TrainSet ts, const FitnessOptions& opt)
  : opt_(opt), spec_(spec), ts_(std::move(ts))
{
  setup();
}

Fitness::~Fitness()
{
  plans_.clear();
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Fitness::device_geometry(std::vector<float4>& bond, std::vector<float4>& tri,
                              std::vector<float4>& rec) const
{
  bond.resize(ds_.bond.size());
  tri.resize(ds_.tri.size());
  rec.resize(ds_.rec.size());
  ds_.bond.download(bond.data(), bond.size(), stream_);
  ds_.tri.download(tri.data(), tri.size(), stream_);
  ds_.rec.download(rec.data(), rec.size(), stream_);
  GPUGA_HIP(hipStreamSynchronize(stream_));
}

void Fitness::setup()
{
  int ndev = 0;
  GPUGA_HIP(hipGetDeviceCount(&ndev));
  if (ndev < 1) fail("no HIP device available");
  GPUGA_HIP(hipGetDevice(&device_));
  hipDeviceProp_t prop;
  GPUGA_HIP(hipGetDeviceProperties(&prop, device_));
  num_cu_ = prop.multiProcessorCount;
  GPUGA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  GPUGA_HIP(hipEventCreate(&ev0_));
  GPUGA_HIP(hipEventCreate(&ev1_));

  const bool tersoff = spec_.type == POT_TERSOFF1 || spec_.type == POT_TERSOFF2;
  if (spec_.type == POT_RI || spec_.type == POT_NN2B_SPECIES)
    for (int t : ts_.type)
      if (t != 0 && t != 1)
        fail(spec_.type == POT_RI ? "RI potential needs atom types 0 and 1"
                                  : "NN2B with the species input needs atom types 0 and 1");
  nl_ = build_neighbors_device(ts_, spec_.cutoff, stream_);
  for (int i = 0; i < ts_.N; ++i)
    for (int s = 0; s < nl_.NN[i]; ++s)
      if (nl_.NL[nl_.rev[nl_.start[i] + s]] != i)
        fail("neighbour list is not symmetric (reverse bond missing)");
  // pair types on the host (the list is there); the geometry tables and the
  // atom records on the device, from the device copy of the list
  const std::vector<int> ptype = bond_types(ts_, nl_, tersoff);

  hipStream_t st = stream_;
  ds_.Na.upload(ts_.Na, st);
  ds_.Na_sum.upload(ts_.Na_sum, st);
  ds_.type.upload(ts_.type, st);
  ds_.nn.upload(nl_.NN, st);
  ds_.bond0.upload(nl_.start, st);
  ds_.nl.upload(nl_.NL.empty() ? std::vector<int>{0} : nl_.NL, st);
  ds_.rev.upload(nl_.rev.empty() ? std::vector<int>{0} : nl_.rev, st);
  ds_.ptype.upload(ptype.empty() ? std::vector<int>{0} : ptype, st);
  DeviceNeighbors dn;
  dn.NN = ds_.nn.get();
  dn.start = ds_.bond0.get();
  dn.NL = ds_.nl.get();
  dn.rev = ds_.rev.get();
  dn.total = nl_.total;
  if (nl_.total == 0) ds_.bond.upload(std::vector<float4>{make_float4(0, 0, 0, 1)}, st);
  else ds_.bond.resize(nl_.total);
  std::vector<int> tri0;
  if (tersoff) {
    long long tri_total = 0;
    tri0 = triplet_offsets(nl_, &tri_total);
    ds_.tri0.upload(tri0, st);
    if (tri_total == 0) ds_.tri.upload(std::vector<float4>(1), st);
    else ds_.tri.resize((size_t)tri_total);
    dn.tri0 = ds_.tri0.get();
    build_geometry_device(ts_, dn, ds_.bond.get(), ds_.tri.get(), st);
    // The population kernel's records leave out the bonds that no
    // individual's taper reaches (exact zeros in every sum).
    const float reach = tersoff_reach(spec_);
    pruned_ = false;
    if (reach < spec_.cutoff && nl_.total > 0) {
      std::vector<float4> b(nl_.total);
      ds_.bond.download(b.data(), b.size(), st);
      GPUGA_HIP(hipStreamSynchronize(st));
      std::vector<float> d(nl_.total);
      for (int k = 0; k < nl_.total; ++k) d[k] = b[k].w;
      pnl_ = prune_neighbors(nl_, d, reach);
      pruned_ = pnl_.total < nl_.total;
      if (!pruned_) pnl_ = NeighborList();
    }
    const NeighborList& pn = population_neighbors();
    const int M = record_slots(pn.max_nn);
    loop_path_ = M == 0;
    if (loop_path_ && pn.max_nn > kLoopMaxNeighbors)
      fail(strfmt("the population kernel supports at most %d neighbours per atom within the "
                  "largest R2 (found %d); use the per-individual backend",
                  kLoopMaxNeighbors, pn.max_nn));
    // tables of the list the population kernel reads: the full ones, or
    // the pruned list's (geometry bit-identical to the full tables' entries
    // for the same bonds)
    DeviceNeighbors pd = dn;
    const int* pptype = ds_.ptype.get();
    const float4 *pbond = ds_.bond.get(), *ptri = ds_.tri.get();
    if (pruned_) {
      long long ptri_total = 0;
      const std::vector<int> t0 = triplet_offsets(pn, &ptri_total);
      ds_.p_nn.upload(pn.NN, st);
      ds_.p_b0.upload(pn.start, st);
      ds_.p_nl.upload(pn.NL.empty() ? std::vector<int>{0} : pn.NL, st);
      ds_.p_rev.upload(pn.rev.empty() ? std::vector<int>{0} : pn.rev, st);
      const std::vector<int> pt = bond_types(ts_, pn, true);
      ds_.p_ptype.upload(pt.empty() ? std::vector<int>{0} : pt, st);
      ds_.p_tri0.upload(t0, st);
      ds_.p_bond.resize(std::max(pn.total, 1));
      ds_.p_tri.resize((size_t)std::max<long long>(ptri_total, 1));
      pd.NN = ds_.p_nn.get();
      pd.start = ds_.p_b0.get();
      pd.NL = ds_.p_nl.get();
      pd.rev = ds_.p_rev.get();
      pd.tri0 = ds_.p_tri0.get();
      pd.total = pn.total;
      build_geometry_device(ts_, pd, ds_.p_bond.get(), ds_.p_tri.get(), st);
      pptype = ds_.p_ptype.get();
      pbond = ds_.p_bond.get();
      ptri = ds_.p_tri.get();
      // (the host vectors above are read before this scope ends)
      GPUGA_HIP(hipStreamSynchronize(st));
    }
    if (M > 0) {
      const size_t RS = 1 + M + (size_t)M * M + M / 2;
      ds_.rec.resize((size_t)ts_.N * RS);
      build_records_device(ts_.N, M, pd, pptype, pbond, ptri, ds_.rec.get(), st);
      ds_.view.rec_nmax = M;
      GPUGA_HIP(hipStreamSynchronize(st));
      // the records hold everything the kernel reads of the pruned list
      for (auto* b : {&ds_.p_nn, &ds_.p_b0, &ds_.p_nl, &ds_.p_rev, &ds_.p_ptype, &ds_.p_tri0})
        b->resize(0);
      ds_.p_bond.resize(0);
      ds_.p_tri.resize(0);
    } else {
      loop_.nn = pd.NN;
      loop_.b0 = pd.start;
      loop_.tri0 = pd.tri0;
      loop_.rev = pd.rev;
      loop_.ptype = pptype;
      loop_.bond = pbond;
      loop_.tri = ptri;
      loop_.max_nn = pn.max_nn;
    }
    ds_.view.bond_force =
      ts_.N_force > 0 ? pn.start[ts_.N_force - 1] + pn.NN[ts_.N_force - 1] : 0;
    dmin_ = dmax_ = 1.0f; // (the NN2B tabulation range; unused here)
  } else {
    ds_.tri0.upload(std::vector<int>(ts_.N, 0), st);
    ds_.tri.upload(std::vector<float4>(1), st);
    build_geometry_device(ts_, dn, ds_.bond.get(), nullptr, st);
    // the pair tables below are built on the host from the bond geometry
    geo_.bond.resize(ds_.bond.size());
    static_assert(sizeof(Vec4) == sizeof(float4), "layout");
    ds_.bond.download(reinterpret_cast<float4*>(geo_.bond.data()), geo_.bond.size(), st);
    GPUGA_HIP(hipStreamSynchronize(st));
    dmin_ = 1e30f;
    dmax_ = 0.0f;
    for (int b = 0; b < nl_.total; ++b) {
      dmin_ = std::min(dmin_, geo_.bond[b].w);
      dmax_ = std::max(dmax_, geo_.bond[b].w);
    }
    if (nl_.total == 0) dmin_ = dmax_ = 1.0f;
  }
  // the uploads above read host vectors of this scope
  GPUGA_HIP(hipStreamSynchronize(st));
  // the pairs a pair potential evaluates for atom i: all its bonds in force
  // configurations (forces need both directions), j > i elsewhere (energy
  // and virial need each pair once); pair_offsets()[i] = the first of atom
  // i's pairs, [N] = their number
  auto evaluates = [&](int i, int b) { return i < ts_.N_force || nl_.NL[b] > i; };
  auto pair_offsets = [&] {
    std::vector<int> off(ts_.N + 1, 0);
    parallel_for(ts_.N, [&](int i) {
      int c = 0;
      for (int b = nl_.start[i]; b < nl_.start[i] + nl_.NN[i]; ++b) c += evaluates(i, b);
      off[i + 1] = c;
    });
    for (int i = 0; i < ts_.N; ++i) off[i + 1] += off[i];
    return off;
  };
  if (spec_.type == POT_RI) {
    // per-bond factors of src/ri.cu:80-101 that do not depend on parameters
    std::vector<float4> tab(std::max(nl_.total, 1));
    parallel_for(ts_.N, [&](int i) {
      for (int b = nl_.start[i]; b < nl_.start[i] + nl_.NN[i]; ++b) {
        const double d = geo_.bond[b].w;
        const double dinv = 1.0 / d;
        const double e = std::erfc(0.2 * d) * dinv;
        const double g = e * dinv + 0.225675833419103 * std::exp(-0.04 * d * d) * dinv;
        tab[b] = make_float4((float)std::pow(dinv, 6), (float)e, (float)g, (float)dinv);
      }
    });
    ds_.ri_tab.upload(tab, st);
    // population form: per atom the pairs it evaluates (all bonds for force
    // configurations, those with j > i for the others), grouped by pair type
    // t_i + t_j.  A force configuration's pair: unit vector and d (one
    // float4); an energy configuration's pair: d and the virial coefficients
    // u_a u_b d (xx, yy, zz, then xy, yz, zx; two float4: the virial then
    // takes one FMA per component).  The header's first word is the atom's
    // first float4.  Of a pair's terms (src/ri.cu:77-110) only the
    // Buckingham exponential and the Morse term are non-linear in the
    // parameters; the others are a parameter times a function of d, so their
    // sums over a configuration (energy, virial) or over an atom's bonds
    // (force) are taken here once, in FP64, and the kernel adds the
    // parameters' combination of them once per configuration or atom.
    const std::vector<int> p0 = pair_offsets();
    auto rows_before = [&](int i) { // float4 rows of the atoms before i
      const long nf = std::min(i, ts_.N_force);
      return (long)p0[nf] + 2 * ((long)p0[i] - p0[nf]);
    };
    const long rows = rows_before(ts_.N);
    if (rows > INT_MAX) fail("rigid-ion pair table too large");
    std::vector<float4> pt((size_t)std::max<long>(rows, 1));
    std::vector<int4> hdr(std::max(ts_.N, 1));
    parallel_for(ts_.N, [&](int i) {
      const int nn = nl_.NN[i], b0 = nl_.start[i];
      const bool force = i < ts_.N_force;
      int cnt[3] = {0, 0, 0};
      const long first = rows_before(i);
      float4* out = &pt[(size_t)first];
      for (int t = 0; t < 3; ++t)
        for (int u = 0; u < nn; ++u) {
          const int b = b0 + u;
          if (ptype[b] != t || !evaluates(i, b)) continue;
          const Vec4 r = geo_.bond[b];
          const double dinv = 1.0 / r.w;
          if (force) {
            *out++ = make_float4((float)(r.x * dinv), (float)(r.y * dinv), (float)(r.z * dinv), r.w);
          } else {
            // u_a u_b d = r_a r_b / d
            *out++ = make_float4(r.w, (float)(r.x * r.x * dinv), (float)(r.y * r.y * dinv),
                                 (float)(r.z * r.z * dinv));
            *out++ = make_float4((float)(r.x * r.y * dinv), (float)(r.y * r.z * dinv),
                                 (float)(r.z * r.x * dinv), 0.0f);
          }
          ++cnt[t];
        }
      hdr[i] = make_int4((int)first, cnt[0], cnt[1], cnt[2]);
    });
    ds_.pair_tab.upload(pt, st);
    ds_.pair_hdr.upload(hdr, st);
    // f_6 = 6 d^-7, f_c = erfc(a d)/d, h = -(d/dr) f_c (a = 0.2, as above)
    auto factors = [](double d, double& d6, double& f6, double& fc, double& h) {
      const double dinv = 1.0 / d;
      d6 = std::pow(dinv, 6);
      f6 = 6.0 * d6 * dinv;
      fc = std::erfc(0.2 * d) * dinv;
      h = fc * dinv + 0.225675833419103 * std::exp(-0.04 * d * d) * dinv;
    };
    std::vector<double> cc((size_t)std::max(ts_.Nc, 1) * 3 * kRiTypeConst, 0.0);
    parallel_for(ts_.Nc, [&](int c) {
      if (c < ts_.Nc_force) return;
      double* out = &cc[(size_t)c * 3 * kRiTypeConst];
      for (int i = ts_.Na_sum[c]; i < ts_.Na_sum[c] + ts_.Na[c]; ++i)
        for (int b = nl_.start[i]; b < nl_.start[i] + nl_.NN[i]; ++b) {
          if (!evaluates(i, b)) continue;
          const Vec4 r = geo_.bond[b];
          const double d = r.w;
          double d6, f6, fc, h;
          factors(d, d6, f6, fc, h);
          const double v[6] = {r.x * r.x / d, r.y * r.y / d, r.z * r.z / d,
                               r.x * r.y / d, r.y * r.z / d, r.z * r.x / d};
          double* o = out + (size_t)ptype[b] * kRiTypeConst;
          o[0] += 1.0;
          o[1] += d;
          o[2] += fc;
          o[3] += d6;
          for (int a = 0; a < 6; ++a) {
            o[4 + a] += v[a];
            o[10 + a] += v[a] * f6;
            o[16 + a] += v[a] * h;
          }
        }
    });
    ds_.ri_cconst.upload(cc, st);
    std::vector<float> fcon((size_t)std::max(ts_.N_force, 1) * kRiAtomConst, 0.0f);
    parallel_for(ts_.N_force, [&](int i) {
      double o[kRiAtomConst] = {};
      for (int b = nl_.start[i]; b < nl_.start[i] + nl_.NN[i]; ++b) {
        const Vec4 r = geo_.bond[b];
        const double d = r.w;
        double d6, f6, fc, h;
        factors(d, d6, f6, fc, h);
        const double u[3] = {r.x / d, r.y / d, r.z / d};
        for (int x = 0; x < 3; ++x) {
          double* w = o + ptype[b] * 9 + x * 3;
          w[0] += u[x] * f6;
          w[1] += u[x];
          w[2] += u[x] * h;
        }
      }
      for (int k = 0; k < kRiAtomConst; ++k) fcon[(size_t)i * kRiAtomConst + k] = (float)o[k];
    });
    ds_.ri_fconst.upload(fcon, st);
  }
  if (spec_.nn2b()) {
    // population form: per atom the pairs it evaluates (all bonds for force
    // configurations, j > i for the others) in order of distance, so that
    // consecutive pairs read neighbouring rows of the tabulated network; per
    // pair (unit vector, d), the cubic Hermite weights on the tabulation
    // grid (h00, h h10, h01, h h11), and (fc, fc', grid cell, species code)
    // with the reference's fixed taper (src/nn2b.cu:80-92)
    const NN2BGrid grid = nn2b_grid_desc();
    const std::vector<int> p0 = pair_offsets();
    // an energy configuration's pair carries its virial coefficients
    // u_a u_b d (xx, yy, zz, xy; yz, zx) in place of (unit vector, d): four
    // float4 instead of three; the header's first word is the atom's first
    // float4
    auto rows_before = [&](int i) {
      const long nf = std::min(i, ts_.N_force);
      return 3 * (long)p0[nf] + 4 * ((long)p0[i] - p0[nf]);
    };
    const long rows = rows_before(ts_.N);
    if (rows > INT_MAX) fail("NN2B pair table too large");
    std::vector<float4> pt((size_t)std::max<long>(rows, 1));
    std::vector<int4> hdr(std::max(ts_.N, 1));
    parallel_for(ts_.N, [&](int i) {
      const int nn = nl_.NN[i], b0 = nl_.start[i];
      const bool force = i < ts_.N_force;
      std::vector<int> order;
      order.reserve(nn);
      for (int u = 0; u < nn; ++u)
        if (evaluates(i, b0 + u)) order.push_back(b0 + u);
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int b) { return geo_.bond[a].w < geo_.bond[b].w; });
      const long first = rows_before(i);
      hdr[i] = make_int4((int)first, (int)order.size(), 0, 0);
      float4* out = &pt[(size_t)first];
      for (int b : order) {
        const Vec4 r = geo_.bond[b];
        const double d = r.w, dinv = 1.0 / d;
        const double u = (d - (double)grid.d0) / (double)grid.h;
        const int g = std::max(0, std::min((int)u, grid.G - 2));
        const double t = u - g, t2 = t * t, t3 = t2 * t, h = grid.h;
        double fc = 0.0, fcp = 0.0;
        if (d < NN2B_R1) {
          fc = 1.0;
        } else if (d < NN2B_R2) {
          const double a = (double)NN2B_PI_FACTOR * (d - NN2B_R1);
          fc = 0.5 * std::cos(a) + 0.5;
          fcp = -0.5 * (double)NN2B_PI_FACTOR * std::sin(a);
        }
        if (force) {
          *out++ = make_float4((float)(r.x * dinv), (float)(r.y * dinv), (float)(r.z * dinv), r.w);
        } else {
          *out++ = make_float4((float)(r.x * r.x * dinv), (float)(r.y * r.y * dinv),
                               (float)(r.z * r.z * dinv), (float)(r.x * r.y * dinv));
          *out++ = make_float4((float)(r.y * r.z * dinv), (float)(r.z * r.x * dinv), 0.0f, 0.0f);
        }
        *out++ = make_float4((float)(2 * t3 - 3 * t2 + 1), (float)(h * (t3 - 2 * t2 + t)),
                             (float)(-2 * t3 + 3 * t2), (float)(h * (t3 - t2)));
        // (the species code t_i + t_j picks the individual's table for it)
        const int code = spec_.type == POT_NN2B_SPECIES ? ptype[b] : 0;
        float4 c = make_float4((float)fc, (float)fcp, 0.0f, 0.0f);
        std::memcpy(&c.z, &g, sizeof g);
        std::memcpy(&c.w, &code, sizeof code);
        *out++ = c;
      }
    });
    ds_.pair_tab.upload(pt, st);
    ds_.pair_hdr.upload(hdr, st);
  }
  ds_.pe_ref.upload(ts_.pe_ref, st);
  ds_.virial_ref.upload(ts_.virial_ref, st);
  ds_.force_ref.upload(ts_.force_ref, st);
  GPUGA_HIP(hipStreamSynchronize(st));

  DevSet& v = ds_.view;
  v.Nc = ts_.Nc;
  v.Nc_force = ts_.Nc_force;
  v.N = ts_.N;
  v.N_force = ts_.N_force;
  v.Na = ds_.Na.get();
  v.Na_sum = ds_.Na_sum.get();
  v.type = ds_.type.get();
  v.nn = ds_.nn.get();
  v.bond0 = ds_.bond0.get();
  v.nl = ds_.nl.get();
  v.rev = ds_.rev.get();
  v.ptype = ds_.ptype.get();
  v.bond = ds_.bond.get();
  v.tri0 = ds_.tri0.get();
  v.tri = ds_.tri.get();
  v.rec = ds_.rec.get();
  v.pair_hdr = ds_.pair_hdr.get();
  v.pair_tab = ds_.pair_tab.get();
  v.ri_cconst = ds_.ri_cconst.get();
  v.ri_fconst = ds_.ri_fconst.get();
  v.pe_ref = ds_.pe_ref.get();
  v.virial_ref = ds_.virial_ref.get();
  v.force_ref = ds_.force_ref.get();

  potential_ = make_potential(spec_, ds_);
  potential_->initialize(ts_.N, ts_.max_Na);
  // this device's fixed workspace of the allocation-free path, looked up
  // here (set-up) so that the path's calls only use the kept address
  ws_static_ = static_workspace();
  d_pmin_.upload(spec_.pmin, st);
  d_pmax_.upload(spec_.pmax, st);
  GPUGA_HIP(hipStreamSynchronize(st));
  d_f_.resize(3 * (size_t)ts_.N);
  d_virial_.resize(6 * (size_t)ts_.N);
  d_pe_.resize(ts_.N);
  d_cerr_.resize(3 * (size_t)ts_.Nc);
  d_fit1_.resize(1);
  d_terms1_.resize(3);
}

NN2BGrid Fitness::nn2b_grid_desc() const
{
  NN2BGrid g;
  g.G = std::max(opt_.nn2b_grid, 16);
  g.n = spec_.num_neurons;
  g.bf16 = opt_.nn2b_bf16;
  g.codes = spec_.pair_codes();
  // one spare cell on each side of the training set's pair-distance range
  const float span = std::max(dmax_ - dmin_, 1e-3f);
  g.h = span / (float)(g.G - 3);
  g.d0 = dmin_ - g.h;
  return g;
}

FitnessNorm Fitness::norm() const
{
  FitnessNorm n;
  n.inv_pss = 1.0 / ts_.potential_square_sum;
  n.inv_vss = 1.0 / ts_.virial_square_sum;
  n.inv_fss = 1.0 / ts_.force_square_sum;
  n.w_energy = spec_.w_energy;
  n.w_stress = spec_.w_stress;
  n.w_force = spec_.w_force;
  return n;
}

size_t Fitness::device_bytes() const
{
  size_t b = ds_.bytes() + d_f_.bytes() + d_virial_.bytes() + d_pe_.bytes() + d_cerr_.bytes();
  for (auto& kv : plans_) {
    const PopPlan& p = *kv.second;
    b += p.items.bytes() + p.part.bytes() + p.err.bytes() + p.bondf.bytes() + p.genes.bytes() +
         p.out.bytes() + p.gpart.bytes() + p.counters.bytes() + p.nn2b_table.bytes();
  }
  return b;
}

// ------------------------------------------------------- population plans

namespace {

long long atom_cost(int type, int nn, bool force_cfg)
{
  if (type == POT_TERSOFF1 || type == POT_TERSOFF2) return (long long)nn * nn + 4LL * nn + 8;
  return (force_cfg ? nn : nn / 2 + 1) + 4;
}

int nn_class(int max_nn)
{
  if (max_nn <= 4) return 0;
  if (max_nn <= 8) return 1;
  return 2;
}

} // namespace

size_t Fitness::plan_bytes(int P) const
{
  const size_t K = (P + WAVE - 1) / WAVE;
  const NeighborList& pn = population_neighbors();
  const size_t B_force = ts_.N_force > 0 ? (size_t)pn.start[ts_.N_force - 1] +
                                             pn.NN[ts_.N_force - 1]
                                         : 0;
  const bool tersoff = spec_.type == POT_TERSOFF1 || spec_.type == POT_TERSOFF2;
  size_t b = K * 3 * ts_.Nc * WAVE * sizeof(double);              // err
  b += (size_t)2 * ts_.Nc * K * 8 * WAVE * sizeof(double);         // part (typical)
  if (tersoff) b += K * (B_force + ts_.N_force) * 3 * WAVE * sizeof(float);       // bondf
  if (spec_.nn2b())
    b += (size_t)3 * spec_.pair_codes() * opt_.nn2b_grid * K * WAVE * sizeof(float);
  return b;
}

LaunchShape Fitness::option_shape() const
{
  LaunchShape sh;
  sh.chunk_cost = opt_.chunk_cost;
  sh.queues = opt_.queues;
  sh.blocks_per_cu = opt_.blocks_per_cu;
  sh.pack = opt_.pack_lanes;
  return sh;
}

void Fitness::build_plan(PopPlan& pl, int P, const LaunchShape& sh)
{
  const int chunk_cost = std::max(sh.chunk_cost, 1);
  const int Nc = ts_.Nc;
  const bool tersoff = spec_.type == POT_TERSOFF1 || spec_.type == POT_TERSOFF2;
  // neighbour counts of the list the kernel reads (Tersoff: the pruned one)
  const NeighborList& pn = tersoff ? population_neighbors() : nl_;
  pl.P = P;
  // the looped Tersoff kernel: one wave per workgroup (its LDS rows)
  pl.block = tersoff && loop_path_ ? WAVE : 256;
  pl.K = (P + WAVE - 1) / WAVE;
  pl.chunk_cost = chunk_cost;
  const int tail_split = opt_.tail_split > 0 ? opt_.tail_split : 1;
  const int K = pl.K;

  struct Chunk {
    int c, chunk, a0, a1, cls;
    long long cost;
  };
  std::vector<Chunk> chunks;
  std::vector<int> nch(Nc), slot0(Nc);
  std::vector<long long> cfg_total(Nc, 0);
  for (int c = 0; c < Nc; ++c)
    for (int i = ts_.Na_sum[c]; i < ts_.Na_sum[c] + ts_.Na[c]; ++i)
      cfg_total[c] += atom_cost(spec_.type, pn.NN[i], c < ts_.Nc_force);
  // chunks of configuration c for an item cost of `cost`
  auto chunks_of = [&](int c, long long cost) {
    const long long total = cfg_total[c];
    cost = std::max<long long>(cost, 1);
    int n = (int)std::max<long long>(1, (total + cost - 1) / cost);
    // A configuration smaller than one item comes at the end of its queue
    // (largest first), where the launch drains: it is cut into pieces of
    // 1/tail_split of an item, so that the last items are short.  (Not the
    // three-body force configurations: they come first.)
    if (total < cost && !(tersoff && c < ts_.Nc_force)) {
      const long long piece = std::max<long long>(1, cost / std::max(tail_split, 1));
      n = (int)std::max<long long>(1, (total + piece - 1) / piece);
    }
    return std::min(n, ts_.Na[c]);
  };
  // A three-body force configuration runs two dependent steps (its chunks,
  // then their force assembly), an energy configuration one.  When every item
  // has a SIMD of its own (a launch too small to fill the device, so its time
  // is the longest chain rather than the total work), the force
  // configurations' chunks are cut at half the item cost, so that both
  // chains end together.  (Cut at a third, the waves of a launch of this
  // kind went from 768 to 984 and every atom took longer: 1.03 against 0.87
  // us per energy atom.)
  long long force_cost = chunk_cost;
  if (tersoff && ts_.Nc_force > 0) {
    // (a force configuration of several chunks has an assembly item per
    // chunk; 4 SIMDs per CU)
    long long n_items = 0;
    for (int c = 0; c < Nc; ++c) {
      const bool fc = c < ts_.Nc_force;
      const int n = chunks_of(c, fc ? chunk_cost / 2 : chunk_cost);
      n_items += (long long)n * (fc && n > 1 ? 2 : 1);
    }
    if (n_items * K <= 4LL * num_cu_) force_cost = std::max(chunk_cost / 2, 1);
  }
  int slots = 0;
  for (int c = 0; c < Nc; ++c) {
    const int a0 = ts_.Na_sum[c], a1 = a0 + ts_.Na[c];
    const bool fc = c < ts_.Nc_force;
    const long long total = cfg_total[c];
    const int n = chunks_of(c, fc ? force_cost : chunk_cost);
    nch[c] = n;
    slot0[c] = slots;
    slots += n;
    long long acc = 0;
    int start = a0, k = 0, mx = 0;
    long long cacc = 0;
    for (int i = a0; i < a1; ++i) {
      const long long ci = atom_cost(spec_.type, pn.NN[i], fc);
      acc += ci;
      cacc += ci;
      mx = std::max(mx, pn.NN[i]);
      // cut at the k-th cost quantile, or when the atoms left are exactly
      // the chunks still to make
      const int left_atoms = a1 - (i + 1), left_chunks = n - 1 - k;
      const bool must = left_chunks > 0 && left_atoms == left_chunks;
      const bool want = left_chunks > 0 && left_atoms >= left_chunks && acc * n >= total * (k + 1);
      if (must || want || i == a1 - 1) {
        chunks.push_back(Chunk{c, k, start, i + 1, nn_class(mx), cacc});
        start = i + 1;
        ++k;
        mx = 0;
        cacc = 0;
      }
    }
    if (k != n) fail("internal: chunk split mismatch");
  }
  std::stable_sort(chunks.begin(), chunks.end(),
                   [](const Chunk& a, const Chunk& b) { return a.cost > b.cost; });
  // Three-body force configurations first: their force-assembly items then
  // run while the energy configurations' chunks are still being served, and
  // the end of the launch is the smallest energy chunks, which wait for
  // nothing (largest first within each part).  Ordered by size alone, small
  // force configurations came last and their assembly items formed a tail:
  // 20% of the launch with fewer and fewer waves busy.
  if (tersoff)
    std::stable_partition(chunks.begin(), chunks.end(),
                          [&](const Chunk& a) { return a.c < ts_.Nc_force; });
  // A last lane block of at most 16 individuals runs packed: G >= 4 groups
  // of W lanes, each group on every G-th atom of the chunk.  (A packed atom
  // step costs several uniform ones, since its geometry is per lane: with
  // G = 2, P = 216, the launch took 0.607 ms packed against 0.521 ms
  // unpacked; with G = 4, P = 208, 0.458 against 0.523 ms.)
  int G_last = 1;
  if (tersoff && sh.pack && !loop_path_) {
    const int Pr = P - (K - 1) * WAVE;
    int W = 1;
    while (W < Pr) W <<= 1;
    if (W <= 16) G_last = WAVE / W;
  }
  pl.groups_last = G_last;
  struct Costed {
    WorkItem w;
    long long cost;
  };
  std::vector<Costed> eval;
  eval.reserve(chunks.size() * K);
  for (const Chunk& ch : chunks)
    for (int k = 0; k < K; ++k) {
      const int G = k == K - 1 ? G_last : 1;
      const int n = ch.a1 - ch.a0, steps = (n + G - 1) / G;
      // (with the assembly of a one-chunk three-body force configuration)
      const long long asm_cost = tersoff && ch.c < ts_.Nc_force && nch[ch.c] == 1 ? 8LL * steps : 0;
      eval.push_back(Costed{WorkItem{ch.c, k, ch.chunk, ch.a0, ch.a1, ch.cls, 0, G},
                            ch.cost * steps / n + asm_cost});
    }
  // three-body force configurations of several chunks: one force-assembly
  // item per chunk and lane block (a configuration of one chunk is
  // assembled by its phase-0 item)
  std::vector<Costed> assemble;
  if (tersoff)
    for (const Chunk& ch : chunks)
      if (ch.c < ts_.Nc_force && nch[ch.c] > 1)
        for (int k = 0; k < K; ++k) {
          const int G = k == K - 1 ? G_last : 1;
          const int n = ch.a1 - ch.a0, steps = (n + G - 1) / G;
          assemble.push_back(
            Costed{WorkItem{ch.c, k, ch.chunk, ch.a0, ch.a1, ch.cls, 1, G}, 8LL * steps});
        }
  pl.n_items = (int)(eval.size() + assemble.size());
  const int waves_per_block = pl.block / WAVE;
  {
    int occ = sh.blocks_per_cu;
    if (occ <= 0) {
      switch (spec_.type) {
      case POT_RI: occ = ri_population_occupancy(pl.block); break;
      case POT_NN2B:
      case POT_NN2B_SPECIES: occ = nn2b_population_occupancy(pl.block); break;
      default:
        occ = loop_path_ ? tersoff_loop_occupancy(spec_.type == POT_TERSOFF2, pn.max_nn, pl.block)
                         : tersoff_population_occupancy(nn_class(pn.max_nn),
                                                        spec_.type == POT_TERSOFF2, G_last > 1,
                                                        pl.block);
        break;
      }
      occ = std::max(occ, 1);
    }
    const int need = (pl.n_items + waves_per_block - 1) / waves_per_block;
    pl.grid = std::max(1, std::min(need, num_cu_ * occ));
    pl.occ = occ;
  }

  // Queues, nq / 8 per XCD (queue q is served first by XCD q / (nq / 8);
  // four per XCD unless the shape says otherwise): whole configurations,
  // balanced by cost (largest configuration to the least loaded queue).
  // Within a queue, phase-0 items largest chunk first, the lane blocks of a
  // chunk together; a force-assembly item follows the last phase-0 item of
  // its configuration by about one item per resident wave of the queue.
  // (at most 64: the kernel keeps one bit per queue in its closed mask)
  const int nq = std::max(1, std::min({sh.queues > 0 ? sh.queues : kQueues,
                                       pl.grid * waves_per_block, 64}));
  std::vector<long long> cfg_cost(Nc, 0);
  for (const Costed& e : eval) cfg_cost[e.w.c] += e.cost;
  for (const Costed& e : assemble) cfg_cost[e.w.c] += e.cost;
  std::vector<int> by_cost(Nc), cfg_q(Nc, 0);
  std::iota(by_cost.begin(), by_cost.end(), 0);
  std::stable_sort(by_cost.begin(), by_cost.end(),
                   [&](int a, int b) { return cfg_cost[a] > cfg_cost[b]; });
  std::vector<long long> load(nq, 0);
  for (int c : by_cost) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    cfg_q[c] = q;
    load[q] += cfg_cost[c];
  }
  // eval is in chunk order (largest chunk first, its lane blocks together):
  // a packed item stays beside its siblings, so the packed block's
  // force-assembly items do not all wait at the end of the launch
  const long long slack = std::max(1, pl.grid * waves_per_block / nq);
  std::vector<WorkItem> items;
  items.reserve(pl.n_items);
  std::vector<int> qstart(nq + 1, 0);
  for (int q = 0; q < nq; ++q) {
    qstart[q] = (int)items.size();
    // (position key, tie, item): phase-0 items at their rank in the queue,
    // assembly items at rank of their configuration's last phase-0 item +
    // slack (after it in any case)
    struct Keyed {
      long long key;
      int tie;
      WorkItem w;
    };
    std::vector<Keyed> qi;
    std::vector<long long> last(Nc, -1);
    long long pos = 0;
    for (const Costed& e : eval)
      if (cfg_q[e.w.c] == q) {
        qi.push_back(Keyed{pos, 0, e.w});
        last[e.w.c] = pos++;
      }
    int tie = 1;
    for (const Costed& e : assemble)
      if (cfg_q[e.w.c] == q) qi.push_back(Keyed{last[e.w.c] + slack, tie++, e.w});
    std::stable_sort(qi.begin(), qi.end(), [](const Keyed& a, const Keyed& b) {
      return a.key != b.key ? a.key < b.key : a.tie < b.tie;
    });
    for (const Keyed& x : qi) items.push_back(x.w);
  }
  qstart[nq] = (int)items.size();
  pl.nq = nq;

  hipStream_t st = stream_;
  pl.items.upload(items, st);
  pl.qstart.upload(qstart, st);
  // When the grid has a wave for every item, the first wave_slots are spread
  // over the queues in proportion to their lengths (wave w starts at the
  // queue that holds item w), so that each wave's first fetch finds an item.
  // Queues are balanced by cost, not count: with home queues by XCD alone, a
  // queue longer than its share of waves left its last items to waves that
  // had gone through the empty queues first (about 2.6 us per queue tried),
  // and the phase-0 items of small force configurations started up to 7 us
  // late.
  // (kept until the stream synchronisation below: the copy reads it)
  std::vector<int> home;
  int n_home = 0;
  if ((long long)pl.n_items <= (long long)pl.grid * waves_per_block) {
    home.resize(pl.n_items);
    for (int q = 0; q < nq; ++q)
      for (int i = qstart[q]; i < qstart[q + 1]; ++i) home[i] = q;
    pl.home.upload(home, st);
    n_home = pl.n_items;
  }
  pl.cfg_nchunks.upload(nch, st);
  pl.cfg_slot0.upload(slot0, st);
  pl.part.resize((size_t)std::max(slots, 1) * K * 8 * WAVE);
  pl.err.resize((size_t)K * 3 * Nc * WAVE);
  // results in one buffer, so one copy returns them: terms [P][3] (FP64),
  // fitness [P] (FP32), status word
  const size_t out_doubles = 3 * (size_t)P + ((size_t)P + 1 + 1) / 2;
  pl.out.resize(out_doubles);
  pl.out.zero(st);
  pl.genes.resize((size_t)P * spec_.V);
  const int B_force =
    ts_.N_force > 0 ? pn.start[ts_.N_force - 1] + pn.NN[ts_.N_force - 1] : 0;
  pl.bondf.resize(tersoff ? std::max<size_t>((size_t)K * (B_force + ts_.N_force) * 3 * WAVE, 1) : 1);
  if (spec_.nn2b())
    pl.nn2b_table.resize((size_t)spec_.pair_codes() * K * nn2b_grid_desc().G * 3 * WAVE);
  const int cfg_group = config_group(Nc);
  const int groups = (Nc + cfg_group - 1) / cfg_group;
  pl.gpart.resize((size_t)K * groups * 3 * WAVE);
  // counters: cnt [Nc*K] | cnt2 [Nc*K] | grp_done [K*groups] | blk_done [K] |
  // (pad to 4 KiB) fetch [nq, kFetchStride apart] | done
  const size_t fetch0 = ((size_t)2 * Nc * K + (size_t)K * groups + K + 1023) / 1024 * 1024;
  // ... then done (one word) and, 8-byte aligned in the next 256 bytes, the
  // closed-queue mask
  const size_t closed0 = fetch0 + (size_t)nq * kFetchStride + 64;
  pl.counters.resize(closed0 + 2);
  pl.counters.zero(st);
  pl.h_genes.resize((size_t)P * spec_.V);
  pl.h_out.resize(out_doubles);
  std::memset(pl.h_out.get(), 0, sizeof(double) * out_doubles);

  PopCtl& c = pl.ctl;
  c.items = pl.items.get();
  c.n_items = pl.n_items;
  c.nq = nq;
  c.qstart = pl.qstart.get();
  c.home = n_home > 0 ? pl.home.get() : nullptr;
  c.n_home = n_home;
  c.K = K;
  c.P = P;
  c.V = spec_.V;
  c.cfg_nchunks = pl.cfg_nchunks.get();
  c.cfg_slot0 = pl.cfg_slot0.get();
  c.genes = pl.genes.get();
  c.pmin = d_pmin_.get();
  c.pmax = d_pmax_.get();
  c.part = pl.part.get();
  c.bondf = pl.bondf.get();
  c.B_force = B_force + ts_.N_force; // slots per lane block: bonds, then atoms
  c.err = pl.err.get();
  unsigned* ctr = pl.counters.get();
  c.cnt = ctr;
  c.cnt2 = ctr + (size_t)Nc * K;
  c.grp_done = ctr + (size_t)2 * Nc * K;
  c.blk_done = c.grp_done + (size_t)K * groups;
  c.fetch = ctr + fetch0;
  c.done = c.fetch + (size_t)nq * kFetchStride;
  c.closed = reinterpret_cast<unsigned long long*>(ctr + closed0);
  c.terms = pl.out.get();
  c.fitness = reinterpret_cast<float*>(c.terms + 3 * (size_t)P);
  c.status = reinterpret_cast<unsigned*>(c.fitness + P);
  c.gpart = pl.gpart.get();
  c.cfg_group = cfg_group;
  c.spin_limit = spin_limit_;
  c.norm = norm();
  GPUGA_HIP(hipStreamSynchronize(st));
}

} // namespace gpuga

namespace gpuga {

PopPlan& Fitness::plan(int P)
{
  auto it = plans_.find(P);
  if (it != plans_.end()) return *it->second;
  auto pl = std::make_unique<PopPlan>();
  build_plan(*pl, P, launch_shape(P));
  PopPlan& ref = *pl;
  plans_[P] = std::move(pl);
  return ref;
}

void Fitness::launch_population(PopPlan& pl, hipStream_t st, const float* genes, int grid)
{
  PopCtl ctl = pl.ctl;
  if (genes) ctl.genes = genes;
  // (a persistent grid of any size drains the same queues; the results do
  // not depend on it)
  const int G = grid > 0 ? grid : pl.grid;
  switch (spec_.type) {
  case POT_RI: launch_ri_population(ctl, ds_.view, ds_.ri_tab.get(), G, pl.block, st); break;
  case POT_NN2B:
  case POT_NN2B_SPECIES: {
    NN2BGrid g = nn2b_grid_desc();
    launch_nn2b_tabulate(ctl, g, pl.nn2b_table.get(), st);
    launch_nn2b_population(ctl, ds_.view, g, pl.nn2b_table.get(), G, pl.block, st);
    break;
  }
  default: {
    if (loop_path_) {
      launch_tersoff_loop_population(ctl, ds_.view, loop_, G, pl.block, st);
      break;
    }
    // the instantiation with the packed path only when the plan has packed
    // items; the one for lone waves when there is at most a workgroup per CU
    launch_tersoff_population(ctl, ds_.view, nn_class(population_neighbors().max_nn), G,
                              pl.block, pl.groups_last > 1, G <= num_cu_, st);
    break;
  }
  }
}

void Fitness::check_genes(int P, const float* genes) const
{
  // the pruned records leave out bonds beyond R2 at its upper bound, which
  // bounds R2 for genes in [0, 1] only
  if (!pruned_) return;
  for (int p = 0; p < P; ++p)
    if (!(genes[(size_t)p * spec_.V + spec_.V - 1] <= 1.0f))
      fail("genes must lie in [0, 1] (the outer cut-off radius gene exceeds 1)");
}

void Fitness::compute_async(int P, const float* genes, int share)
{
  if (P < 1) fail("population size must be positive");
  check_genes(P, genes);
  PopPlan& pl = plan(P);
  hipStream_t st = stream_;
  std::memcpy(pl.h_genes.get(), genes, sizeof(float) * (size_t)P * spec_.V);
  GPUGA_HIP(hipMemcpyAsync(pl.genes.get(), pl.h_genes.get(), sizeof(float) * (size_t)P * spec_.V,
                           hipMemcpyHostToDevice, st));
  if (timing_) GPUGA_HIP(hipEventRecord(ev0_, st));
  launch_generation(pl, st, share);
  if (timing_) GPUGA_HIP(hipEventRecord(ev1_, st));
  GPUGA_HIP(hipMemcpyAsync(pl.h_out.get(), pl.out.get(),
                           sizeof(double) * 3 * P + sizeof(float) * P + sizeof(unsigned),
                           hipMemcpyDeviceToHost, st));
  pending_P_ = P;
}

namespace {

// The launches of `body` on st as a graph, instantiated.  A failure inside
// ends the capture first, so the stream is usable again.
template <class F>
void captured(hipStream_t st, F&& body, hipGraph_t* graph, hipGraphExec_t* exec)
{
  GPUGA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  GPUGA_HIP(hipStreamEndCapture(st, graph));
  GPUGA_HIP(hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
  // the graph's upload to the device, now: left to its first launch, it
  // held that launch back by up to 2 ms with the device idle (an 8-
  // generation graph of the 384-individual Tersoff workload)
  GPUGA_HIP(hipGraphUpload(*exec, st));
}

} // namespace

void Fitness::launch_generation(PopPlan& pl, hipStream_t st, int share)
{
  // share 1, 2, 4 or 8 (rounded up, at most 8): 1/share of the grid
  int j = 0;
  while (j < PopPlan::kShares && (1 << j) < share) ++j;
  const int grid = std::max(1, (pl.grid + (1 << j) - 1) >> j);
  if (!opt_.use_graph) {
    launch_population(pl, st, nullptr, grid);
    return;
  }
  if (j == 0) {
    capture(pl, st);
    GPUGA_HIP(hipGraphLaunch(pl.exec, st));
    return;
  }
  if (!pl.share_exec[j - 1])
    captured(st, [&] { launch_population(pl, st, nullptr, grid); }, &pl.share_graph[j - 1],
             &pl.share_exec[j - 1]);
  GPUGA_HIP(hipGraphLaunch(pl.share_exec[j - 1], st));
}

void Fitness::capture(PopPlan& pl, hipStream_t st)
{
  if (pl.exec) return;
  captured(st, [&] { launch_population(pl, st); }, &pl.graph, &pl.exec);
}

void Fitness::prepare(int P)
{
  if (opt_.backend == 1) return;
  PopPlan& pl = plan(P);
  if (opt_.use_graph) capture(pl, stream_);
}

static_assert((1 << (PopPlan::kGaGraphs - 1)) >= kGaMaxBatch, "a batch fits the largest graph");

void Fitness::capture_generations(PopPlan& pl, int k, int n, hipGraph_t* graph,
                                  hipGraphExec_t* exec)
{
  hipStream_t st = stream_;
  captured(st, [&] {
    for (int i = 0; i < n; ++i) {
      const int b = (k + i) & 1;
      launch_population(pl, st, pl.dga->genes(b));
      pl.dga->step(b, pl.ctl.fitness, st);
    }
  }, graph, exec);
}

void Fitness::run_ga(GA& ga, int first_generation, int generations, float* best, float* elite)
{
  // GA::compute (src/ga.cu:73-92) for `generations` generations with the
  // population on the device: each generation is the population launch on
  // one population buffer followed by the device GA step into the other;
  // the GA's random decisions are drawn on the host a batch ahead (they do
  // not depend on the fitness) and nothing returns to the host until the end
  const int P = ga.params().population_size, V = spec_.V;
  if (ga.V() != V) fail("the GA's parameter count differs from the potential's");
  if (first_generation < 0) fail("generation numbers start at 0");
  if (generations < 1) return;
  check_genes(P, ga.population().data());
  PopPlan& pl = plan(P);
  hipStream_t st = stream_;
  // the GA as it was before this call: staging draws advances its random
  // stream, so a failed run restores it
  const GA before = ga;
  if (!pl.dga || !pl.dga->matches(ga)) {
    pl.drop_ga_graphs();
    pl.dga = std::make_unique<DeviceGA>(ga, pl.genes.get(), 1024, st, opt_.ga_batch);
  }
  DeviceGA& dga = *pl.dga;
  dga.upload(ga, first_generation, st);
  // batches double from one generation up to the staging capacity, so the
  // device starts after one generation's draws and the host then draws each
  // batch while the device runs the one before
  const int B = dga.batch();
  // Every graph a batch can use (runs of 2^j generations on either
  // population buffer) is captured before the first launch, on the first
  // call: captured on first use, they cost the call that met them about
  // 1-3 ms of host time with the device idle.
  if (opt_.use_graph)
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < PopPlan::kGaGraphs && (1 << j) <= B; ++j)
        if (!pl.ga_exec[k][j]) capture_generations(pl, k, 1 << j, &pl.ga_graph[k][j], &pl.ga_exec[k][j]);
  const int H = dga.history_cap();
  std::vector<float> h((size_t)H * (1 + V));
  int g = 0; // generations queued
  while (g < generations) {
    const int n = std::min(H, generations - g);
    if (g > 0) dga.restart_history(first_generation + g, st);
    if (timing_) GPUGA_HIP(hipEventRecord(ev0_, st));
    for (int b = 0, ramp = 1, nb = 0; b < n; b += nb, ramp = std::min(2 * ramp, B)) {
      nb = std::min({ramp, B, n - b});
      dga.stage(ga, first_generation + g + b, nb, st);
      // the batch as runs of 2^j generations, each one graph (no launch
      // gap between the generations inside a graph)
      for (int i = 0; i < nb;) {
        const int k = (g + b + i) & 1;
        if (!opt_.use_graph) {
          launch_population(pl, st, dga.genes(k));
          dga.step(k, pl.ctl.fitness, st);
          ++i;
          continue;
        }
        int j = 0;
        while (j + 1 < PopPlan::kGaGraphs && (2 << j) <= nb - i) ++j;
        if (!pl.ga_exec[k][j]) capture_generations(pl, k, 1 << j, &pl.ga_graph[k][j], &pl.ga_exec[k][j]);
        GPUGA_HIP(hipGraphLaunch(pl.ga_exec[k][j], st));
        i += 1 << j;
      }
    }
    if (timing_) GPUGA_HIP(hipEventRecord(ev1_, st));
    unsigned status = 0;
    GPUGA_HIP(hipMemcpyAsync(&status, pl.ctl.status, sizeof status, hipMemcpyDeviceToHost, st));
    GPUGA_HIP(hipStreamSynchronize(st));
    if (timing_) {
      float ms = 0.0f;
      GPUGA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
      kernel_ms_ += ms;
      kernel_launches_ += n;
    }
    if (status != 0u) {
      // the device population has advanced on invalid fitness values: drop
      // the device GA, and the GA handle returns to its state before this call
      ga = before;
      pl.dga.reset();
      pl.drop_ga_graphs();
      discard_launch(pl);
    }
    dga.history(h.data(), n, st);
    for (int i = 0; i < n; ++i) {
      if (best) best[g + i] = h[(size_t)i * (1 + V)];
      if (elite)
        std::memcpy(elite + (size_t)(g + i) * V, &h[(size_t)i * (1 + V) + 1], sizeof(float) * V);
    }
    g += n;
  }
  dga.download(ga, generations & 1, st);
}

void Fitness::discard_launch(PopPlan& pl)
{
  // A dependency wait of the population kernel ran out: this launch's results
  // are not valid.  A late phase-0 item may also have counted its chunk after
  // the assembly items returned the counter to zero, so every counter and the
  // result block (status word included) start from zero again before the
  // plan is used next (the launch has ended: the stream is synchronised).
  hipStream_t st = stream_;
  GPUGA_HIP(hipStreamSynchronize(st));
  pl.counters.zero(st);
  pl.out.zero(st);
  GPUGA_HIP(hipStreamSynchronize(st));
  fail("population kernel: force-assembly wait timed out; results discarded");
}

void Fitness::set_spin_limit(int spins)
{
  spin_limit_ = spins;
  for (auto& kv : plans_) {
    PopPlan& pl = *kv.second;
    pl.ctl.spin_limit = spins;
    // the captured launches hold the old control block
    pl.drop_ga_graphs();
    pl.drop_share_graphs();
    if (pl.exec) (void)hipGraphExecDestroy(pl.exec);
    if (pl.graph) (void)hipGraphDestroy(pl.graph);
    pl.exec = nullptr;
    pl.graph = nullptr;
  }
}

void Fitness::compute_wait(float* fitness, double* terms)
{
  const int P = pending_P_;
  if (P < 1) fail("no generation in flight");
  GPUGA_HIP(hipStreamSynchronize(stream_));
  PopPlan& pl = *plans_.at(P);
  if (timing_) {
    float ms = 0.0f;
    GPUGA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    kernel_ms_ += ms;
    ++kernel_launches_;
  }
  pending_P_ = 0;
  const double* h_terms = pl.h_out.get();
  const float* h_fitness = reinterpret_cast<const float*>(h_terms + 3 * (size_t)P);
  const unsigned h_status = *reinterpret_cast<const unsigned*>(h_fitness + P);
  if (h_status != 0u) discard_launch(pl);
  if (fitness) std::memcpy(fitness, h_fitness, sizeof(float) * P);
  if (terms) std::memcpy(terms, h_terms, sizeof(double) * 3 * P);
}

void Fitness::compute(int P, const float* genes, float* fitness, double* terms)
{
  if (opt_.backend == 1) {
    compute_plain(P, genes, fitness, terms);
    return;
  }
  compute_async(P, genes);
  compute_wait(fitness, terms);
}

void Fitness::compute_plain(int P, const float* genes, float* fitness, double* terms)
{
  plain_loop(P, genes, fitness, terms,
             PlainBuffers{nullptr, d_f_.get(), d_virial_.get(), d_pe_.get(), d_fit1_.get(),
                          d_cerr_.get(), d_terms1_.get()});
}

size_t Fitness::carve_plain(char* base, PlainBuffers* b) const
{
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) / 256 * 256;
    char* p = base ? base + off : nullptr;
    off += bytes;
    return p;
  };
  const size_t N = ts_.N, Nc = ts_.Nc;
  PlainBuffers x;
  const size_t ws = potential_->workspace_floats();
  x.ws = ws ? reinterpret_cast<float*>(take(ws * sizeof(float))) : nullptr;
  x.f = reinterpret_cast<float*>(take(3 * N * sizeof(float)));
  x.virial = reinterpret_cast<float*>(take(6 * N * sizeof(float)));
  x.pe = reinterpret_cast<float*>(take(N * sizeof(float)));
  x.fit1 = reinterpret_cast<float*>(take(sizeof(float)));
  x.cerr = reinterpret_cast<double*>(take(3 * Nc * sizeof(double)));
  x.terms1 = reinterpret_cast<double*>(take(3 * sizeof(double)));
  if (b) *b = x;
  return off;
}

size_t Fitness::noalloc_bytes() const { return carve_plain(nullptr, nullptr); }

void Fitness::compute_plain_noalloc(int P, const float* genes, float* fitness, double* terms)
{
  const size_t need = noalloc_bytes();
  if (need > kStaticWorkspaceBytes)
    fail(strfmt("the allocation-free path needs %zu bytes of device workspace for this "
                "training set; the fixed workspace holds %zu",
                need, kStaticWorkspaceBytes));
  if (!ws_static_) fail("the fixed device workspace was not set up on this device");
  int dev = 0;
  GPUGA_HIP(hipGetDevice(&dev));
  if (dev != device_) fail("the handle's device is not the current device");
  PlainBuffers b;
  carve_plain(static_cast<char*>(ws_static_), &b);
  plain_loop(P, genes, fitness, terms, b);
}

void Fitness::plain_loop(int P, const float* genes, float* fitness, double* terms,
                         const PlainBuffers& b)
{
  // src/fitness.cu:316-333: one parameter set at a time
  std::vector<float> params(spec_.V);
  const FitnessNorm nrm = norm();
  hipStream_t st = stream_;
  for (int n = 0; n < P; ++n) {
    denormalise_host(spec_, genes + (size_t)n * spec_.V, params.data());
    potential_->update_potential(params);
    if (timing_) GPUGA_HIP(hipEventRecord(ev0_, st));
    potential_->find_force(ds_.view, b.f, b.virial, b.pe, st, b.ws);
    launch_fitness_terms(ds_.view, nrm, b.f, b.virial, b.pe, b.cerr, b.fit1, b.terms1, st);
    if (timing_) GPUGA_HIP(hipEventRecord(ev1_, st));
    float fit = 0.0f;
    double t[3];
    GPUGA_HIP(hipMemcpyAsync(&fit, b.fit1, sizeof(float), hipMemcpyDeviceToHost, st));
    GPUGA_HIP(hipMemcpyAsync(t, b.terms1, sizeof(t), hipMemcpyDeviceToHost, st));
    GPUGA_HIP(hipStreamSynchronize(st));
    if (timing_) {
      float ms = 0.0f;
      GPUGA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
      kernel_ms_ += ms;
      ++kernel_launches_;
    }
    fitness[n] = fit;
    if (terms) std::memcpy(terms + 3 * (size_t)n, t, sizeof(t));
  }
}

void Fitness::find_force(const float* physical, float* f, float* virial, float* pe)
{
  std::vector<float> params(physical, physical + spec_.V);
  potential_->update_potential(params);
  hipStream_t st = stream_;
  potential_->find_force(ds_.view, d_f_.get(), d_virial_.get(), d_pe_.get(), st);
  GPUGA_HIP(hipGetLastError());
  if (f) d_f_.download(f, 3 * (size_t)ts_.N, st);
  if (virial) d_virial_.download(virial, 6 * (size_t)ts_.N, st);
  if (pe) d_pe_.download(pe, ts_.N, st);
  GPUGA_HIP(hipStreamSynchronize(st));
}

void Fitness::predict(const std::string& dir, const float* elite)
{
  // src/fitness.cu:335-390
  std::vector<float> params(spec_.V);
  denormalise_host(spec_, elite, params.data());
  const size_t N = ts_.N;
  std::vector<float> f(3 * N), w(6 * N), e(N);
  find_force(params.data(), f.data(), w.data(), e.data());
  auto open = [&](const char* name) {
    std::string path = dir + "/" + name;
    FILE* fid = std::fopen(path.c_str(), "w");
    if (!fid) fail(strfmt("Failed to open %s!", path.c_str()));
    return fid;
  };
  FILE* ff = open("force.out");
  for (int n = 0; n < ts_.N_force; ++n)
    std::fprintf(ff, "%g %g %g %g %g %g\n", f[n], f[n + N], f[n + N * 2], ts_.force_ref[n],
                 ts_.force_ref[n + N], ts_.force_ref[n + N * 2]);
  std::fclose(ff);
  // per-configuration sums in FP32, atom by atom, as the reference writes
  // them (src/fitness.cu:335-345): equal per-atom values give the same text
  auto per_config = [&](FILE* fid, const float* data, const float* ref) {
    for (int c = ts_.Nc_force; c < ts_.Nc; ++c) {
      float sum = 0.0f;
      for (int m = 0; m < ts_.Na[c]; ++m) sum += data[ts_.Na_sum[c] + m];
      std::fprintf(fid, "%g %g\n", sum, ref[c]);
    }
  };
  FILE* fe = open("energy.out");
  per_config(fe, e.data(), ts_.pe_ref.data());
  std::fclose(fe);
  FILE* fv = open("virial.out");
  for (int k = 0; k < 6; ++k) per_config(fv, w.data() + N * k, ts_.virial_ref.data() + ts_.Nc * k);
  std::fclose(fv);
}

int Fitness::heuristic_chunk_cost(int P) const
{
  const bool tersoff = spec_.type == POT_TERSOFF1 || spec_.type == POT_TERSOFF2;
  const std::vector<int>& pnn = tersoff ? population_neighbors().NN : nl_.NN;
  long long total = 0, maxcfg = 0;
  for (int c = 0; c < ts_.Nc; ++c) {
    long long t = 0;
    for (int i = ts_.Na_sum[c]; i < ts_.Na_sum[c] + ts_.Na[c]; ++i)
      t += atom_cost(spec_.type, pnn[i], c < ts_.Nc_force);
    total += t;
    maxcfg = std::max(maxcfg, t);
  }
  const long long K = (P + WAVE - 1) / WAVE;
  long long c = total * K / std::max<long long>(1, 16LL * num_cu_);
  c = std::max<long long>(c, 64);
  c = std::min<long long>(c, std::max<long long>(maxcfg, 64));
  return (int)c;
}

size_t Fitness::estimate_bytes(const PotentialSpec& s, const TrainSet& ts, int P,
                               const FitnessOptions& opt)
{
  // neighbour counts on the device (a counting pass of the list build)
  hipStream_t st = nullptr;
  GPUGA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  std::vector<int> nn;
  try {
    nn = count_neighbors_device(ts, s.cutoff, st);
  } catch (...) {
    (void)hipStreamDestroy(st);
    throw;
  }
  GPUGA_HIP(hipStreamDestroy(st));
  long long bonds = 0;
  size_t tri = 0;
  for (int i = 0; i < ts.N; ++i) {
    bonds += nn[i];
    tri += (size_t)nn[i] * nn[i];
  }
  const size_t N = ts.N, Nc = ts.Nc, B = std::max<long long>(bonds, 1);
  const bool tersoff = s.type == POT_TERSOFF1 || s.type == POT_TERSOFF2;
  size_t b = N * sizeof(int) * 5 + Nc * sizeof(int) * 2 + B * (3 * sizeof(int) + 16);
  b += (tersoff ? tri * 16 + B * 4 * 5 : 0) + (s.type == POT_RI ? B * 48 + N * 16 : 0);
  b += s.nn2b() ? B * 48 + N * 16 : 0;
  b += Nc * 7 * sizeof(float) + N * 3 * sizeof(float);  // references
  b += N * 10 * sizeof(float) + Nc * 3 * sizeof(double); // per-atom outputs
  const size_t K = (P + WAVE - 1) / WAVE;
  size_t B_force = 0;
  for (int i = 0; i < ts.N_force; ++i) B_force += nn[i];
  b += K * 3 * Nc * WAVE * sizeof(double) + 2 * Nc * K * 8 * WAVE * sizeof(double);
  if (tersoff) b += K * (B_force + ts.N_force) * 3 * WAVE * sizeof(float);
  if (s.nn2b()) b += (size_t)3 * s.pair_codes() * opt.nn2b_grid * K * WAVE * sizeof(float);
  b += (size_t)P * s.V * sizeof(float) + (size_t)P * 4 * sizeof(double);
  return b;
}

std::vector<double> Fitness::trace(int P, const float* genes)
{
  check_genes(P, genes);
  PopPlan& pl = plan(P);
  hipStream_t st = stream_;
  const size_t n = pl.n_items;
  const size_t F = kTraceFields;
  DeviceBuffer<unsigned long long> buf(std::max<size_t>(F * n, 1));
  buf.zero(st);
  std::vector<WorkItem> items(n);
  pl.items.download(items.data(), n, st);
  GPUGA_HIP(hipMemcpyAsync(pl.genes.get(), genes, sizeof(float) * (size_t)P * spec_.V,
                           hipMemcpyHostToDevice, st));
  pl.ctl.trace = buf.get();
  launch_population(pl, st);
  pl.ctl.trace = nullptr;
  std::vector<unsigned long long> t(F * n);
  buf.download(t.data(), t.size(), st);
  unsigned status = 0;
  GPUGA_HIP(hipMemcpyAsync(&status, pl.ctl.status, sizeof status, hipMemcpyDeviceToHost, st));
  GPUGA_HIP(hipStreamSynchronize(st));
  if (status != 0u) discard_launch(pl);
  int khz = 0;
  GPUGA_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_));
  const double ns_per_tick = 1e6 / std::max(khz, 1);
  unsigned long long t0 = ~0ull;
  for (size_t i = 0; i < n; ++i) t0 = std::min(t0, t[F * i + 1]);
  auto rel = [&](unsigned long long x) { return ((double)x - (double)t0) * ns_per_tick; };
  std::vector<double> rows;
  rows.reserve(12 * n);
  for (size_t i = 0; i < n; ++i) {
    const WorkItem& w = items[i];
    const unsigned long long* x = &t[F * i];
    rows.insert(rows.end(), {(double)w.c, (double)w.k, (double)w.chunk, (double)w.a0,
                             (double)w.a1, (double)w.phase, (double)x[0], rel(x[1]), rel(x[2]),
                             x[3] ? rel(x[3]) : NAN, x[4] ? rel(x[4]) : NAN,
                             x[5] ? rel(x[5]) : NAN});
  }
  return rows;
}

double Fitness::time_plan(int P, const LaunchShape& sh, int reps, int* occ)
{
  PopPlan pl;
  build_plan(pl, P, sh);
  if (occ) *occ = pl.occ;
  std::vector<float> genes((size_t)P * spec_.V);
  for (size_t i = 0; i < genes.size(); ++i) genes[i] = (float)((i * 2654435761u) % 1000) / 1000.0f;
  hipStream_t st = stream_;
  GPUGA_HIP(hipMemcpyAsync(pl.genes.get(), genes.data(), sizeof(float) * genes.size(),
                           hipMemcpyHostToDevice, st));
  launch_population(pl, st); // warm-up
  GPUGA_HIP(hipEventRecord(ev0_, st));
  for (int r = 0; r < reps; ++r) launch_population(pl, st);
  GPUGA_HIP(hipEventRecord(ev1_, st));
  GPUGA_HIP(hipStreamSynchronize(st));
  float ms = 0.0f;
  GPUGA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  return ms / reps;
}

LaunchShape Fitness::launch_shape(int P)
{
  if (opt_.chunk_cost > 0) return option_shape();
  auto t = tuned_.find(P);
  if (t == tuned_.end()) {
    autotune(P);
    t = tuned_.find(P);
  }
  return t->second;
}

void Fitness::plan_shape(int P, int& chunk_cost, int& queues, int& blocks_per_cu,
                         int& groups_last)
{
  if (P < 1) fail("population size must be positive");
  const PopPlan& pl = plan(P);
  chunk_cost = pl.chunk_cost;
  queues = pl.nq;
  blocks_per_cu = pl.occ;
  groups_last = pl.groups_last;
}

int Fitness::autotune(int P)
{
  // Candidate work-item sizes: the whole-configuration cost spread over
  // 1/2 .. 1/32 of the resident waves, and the largest configuration's cost
  // halved step by step (small sets, where latency rather than balance
  // sets the time).  Then, at the best size, the rest of the launch shape
  // that the options leave open: the queue count, one resident workgroup
  // per CU fewer than the occupancy calculator allows, and, when the last
  // lane block qualifies, packing it or not.  Each choice is kept only if
  // its launch is faster.
  const bool tersoff = spec_.type == POT_TERSOFF1 || spec_.type == POT_TERSOFF2;
  const std::vector<int>& pnn = tersoff ? population_neighbors().NN : nl_.NN;
  long long total = 0, maxcfg = 0;
  for (int c = 0; c < ts_.Nc; ++c) {
    long long t = 0;
    for (int i = ts_.Na_sum[c]; i < ts_.Na_sum[c] + ts_.Na[c]; ++i)
      t += atom_cost(spec_.type, pnn[i], c < ts_.Nc_force);
    total += t;
    maxcfg = std::max(maxcfg, t);
  }
  const int K = (P + WAVE - 1) / WAVE;
  const long long waves = (long long)num_cu_ * 4 * 2;
  std::vector<int> cands;
  for (int div : {1, 2, 4, 8, 16}) {
    long long c = total * K / std::max<long long>(1, waves * div / 2);
    c = std::max<long long>(c, 64);
    c = std::min<long long>(c, std::max<long long>(maxcfg, 64));
    cands.push_back((int)c);
  }
  cands.push_back((int)std::max<long long>(maxcfg, 64)); // whole configurations
  for (long long c = maxcfg / 2; c >= 64; c /= 2) cands.push_back((int)c);
  std::sort(cands.begin(), cands.end());
  cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
  LaunchShape best = option_shape();
  best.chunk_cost = cands.back();
  double best_ms = 1e30;
  int best_occ = 0;
  auto consider = [&](const LaunchShape& sh, int reps) {
    int occ = 0;
    const double ms = time_plan(P, sh, reps, &occ);
    if (opt_.verbose)
      std::printf("autotune P=%d chunk_cost=%d queues=%d blocks_per_cu=%d pack=%d: %.4f ms\n", P,
                  sh.chunk_cost, sh.queues, sh.blocks_per_cu, sh.pack, ms);
    if (ms < best_ms) {
      best_ms = ms;
      best = sh;
      best_occ = occ;
    }
  };
  auto with_cost = [&](int c) {
    LaunchShape sh = best;
    sh.chunk_cost = c;
    return sh;
  };
  const LaunchShape base = best;
  for (int c : cands) {
    LaunchShape sh = base;
    sh.chunk_cost = c;
    consider(sh, 3);
  }
  // the neighbourhood of the best (the coarse candidates are a factor of
  // two apart, and launch times differ by about 1% from run to run), the
  // best one timed again alongside
  const int coarse = best.chunk_cost;
  best_ms = 1e30;
  for (double f : {0.6, 0.75, 0.87, 1.0, 1.15, 1.33, 1.6}) {
    const long long c = std::llround(coarse * f);
    if (c >= 64 && c <= std::max<long long>(maxcfg, 64)) consider(with_cost((int)c), 6);
  }
  // queue count (the kernel keeps at most 64)
  if (opt_.queues <= 0) {
    const LaunchShape inc = best;
    for (int nq : {kQueues / 2, kQueues * 2}) {
      LaunchShape sh = inc;
      sh.queues = std::min(nq, 64);
      consider(sh, 6);
    }
  }
  // one resident workgroup per CU fewer (fewer waves sharing a SIMD)
  if (opt_.blocks_per_cu <= 0 && best_occ > 1) {
    LaunchShape sh = best;
    sh.blocks_per_cu = best_occ - 1;
    consider(sh, 6);
  }
  // the packed last lane block against plain lanes
  if (opt_.pack_lanes && tersoff && !loop_path_) {
    const int Pr = P - (K - 1) * WAVE;
    if (Pr <= 16) {
      LaunchShape sh = best;
      sh.pack = 0;
      consider(sh, 6);
    }
  }
  tuned_[P] = best;
  return best.chunk_cost;
}

} // namespace gpuga
