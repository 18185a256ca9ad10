// C ABI (include/gpuga.h): every entry point converts exceptions into a
// status code and a thread-local message.
#include "../../include/gpuga.h"

#include <chrono>
#include <cstring>
#include <string>

#include "convert.h"
#include "fitness.h"
#include "ga.h"
#include "serve.h"

using namespace gpuga;

struct gpuga_fitness {
  std::unique_ptr<Fitness> f;
};
struct gpuga_ga {
  std::unique_ptr<GA> g;
  // device step (gpuga_ga_step_device): its stream, state and fitness buffer
  hipStream_t st = nullptr;
  std::unique_ptr<DeviceGA> dev;
  DeviceBuffer<float> fit;
  ~gpuga_ga()
  {
    dev.reset();
    if (st) (void)hipStreamDestroy(st);
  }
};
struct gpuga_server {
  std::unique_ptr<Server> s;
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F&& fn)
{
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
  } catch (...) {
    g_last_error = "unknown error";
  }
  return 1;
}

FitnessOptions to_options(const gpuga_options* o)
{
  FitnessOptions f;
  if (!o) return f;
  f.backend = o->backend;
  f.chunk_cost = o->chunk_cost;
  f.use_graph = o->use_graph;
  f.nn2b_bf16 = o->nn2b_bf16;
  if (o->nn2b_grid > 0) f.nn2b_grid = o->nn2b_grid;
  f.verbose = o->verbose;
  f.blocks_per_cu = o->blocks_per_cu;
  f.pack_lanes = o->pack_lanes;
  f.queues = o->queues;
  f.ga_batch = o->ga_batch;
  f.tail_split = o->tail_split;
  f.serve_share = o->serve_share;
  return f;
}

void need(const void* p, const char* what)
{
  if (!p) fail(std::string(what) + " is NULL");
}

} // namespace

extern "C" {

const char* gpuga_last_error(void) { return g_last_error.c_str(); }

void gpuga_default_options(gpuga_options* opt)
{
  if (!opt) return;
  FitnessOptions f;
  opt->backend = f.backend;
  opt->chunk_cost = f.chunk_cost;
  opt->use_graph = f.use_graph;
  opt->nn2b_bf16 = f.nn2b_bf16;
  opt->nn2b_grid = f.nn2b_grid;
  opt->verbose = f.verbose;
  opt->blocks_per_cu = f.blocks_per_cu;
  opt->pack_lanes = f.pack_lanes;
  opt->queues = f.queues;
  opt->ga_batch = f.ga_batch;
  opt->tail_split = f.tail_split;
  opt->serve_share = f.serve_share;
}

int gpuga_fitness_create(const char* input_dir, const gpuga_options* opt, gpuga_fitness** out)
{
  return guarded([&] {
    need(input_dir, "input_dir");
    need(out, "out");
    auto h = new gpuga_fitness;
    try {
      h->f = std::make_unique<Fitness>(std::string(input_dir), to_options(opt));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int gpuga_fitness_create_arrays(int potential_type, int num_neurons, float cutoff,
                                float weight_force, float weight_energy, float weight_virial,
                                const float* pmin, const float* pmax, int Nc, int Nc_force,
                                const int* Na, const int* type, const float* box9, const float* r,
                                const float* pe_ref, const float* virial_ref,
                                const float* force_ref, const gpuga_options* opt,
                                gpuga_fitness** out)
{
  return guarded([&] {
    need(out, "out");
    need(Na, "Na");
    need(type, "type");
    need(box9, "box9");
    need(r, "r");
    PotentialSpec s;
    s.type = potential_type;
    s.num_neurons = num_neurons;
    switch (potential_type) {
    case POT_RI: s.V = 15; break;
    case POT_TERSOFF1: s.V = 9; break;
    case POT_TERSOFF2: s.V = 13; break;
    case POT_NN2B:
    case POT_NN2B_SPECIES:
      if (num_neurons < 1) fail("num_neurons should >= 1.");
      s.V = num_neurons * (num_neurons + 4) + 1 +
            (potential_type == POT_NN2B_SPECIES ? num_neurons : 0);
      break;
    default: fail("unsupported potential type.");
    }
    s.cutoff = cutoff;
    s.w_force = weight_force;
    s.w_energy = weight_energy;
    s.w_stress = weight_virial;
    if (s.w_force < 0 || s.w_energy < 0 || s.w_stress < 0) fail("weights should >= 0");
    s.pmin.resize(s.V);
    s.pmax.resize(s.V);
    s.names.resize(s.V);
    for (int v = 0; v < s.V; ++v) {
      if (s.nn2b()) {
        s.pmin[v] = -3.0f;
        s.pmax[v] = 3.0f;
      } else {
        need(pmin, "pmin");
        need(pmax, "pmax");
        s.pmin[v] = pmin[v];
        s.pmax[v] = pmax[v];
      }
    }
    TrainSet ts = make_train(Nc, Nc_force, Na, type, box9, r, pe_ref, virial_ref, force_ref);
    auto h = new gpuga_fitness;
    try {
      h->f = std::make_unique<Fitness>(s, std::move(ts), to_options(opt));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

void gpuga_fitness_destroy(gpuga_fitness* h) { delete h; }

int gpuga_fitness_num_variables(const gpuga_fitness* h)
{
  return h && h->f ? h->f->number_of_variables() : -1;
}

int gpuga_fitness_bounds(const gpuga_fitness* h, float* pmin, float* pmax)
{
  return guarded([&] {
    need(h, "handle");
    const PotentialSpec& s = h->f->spec();
    if (pmin) std::memcpy(pmin, s.pmin.data(), sizeof(float) * s.V);
    if (pmax) std::memcpy(pmax, s.pmax.data(), sizeof(float) * s.V);
  });
}

int gpuga_fitness_info(const gpuga_fitness* h, int* out)
{
  return guarded([&] {
    need(h, "handle");
    need(out, "out");
    const TrainSet& t = h->f->trainset();
    const PotentialSpec& s = h->f->spec();
    out[0] = s.type;
    out[1] = t.Nc;
    out[2] = t.Nc_force;
    out[3] = t.N;
    out[4] = t.N_force;
    out[5] = t.max_Na;
    out[6] = h->f->neighbors().total;
    out[7] = h->f->neighbors().max_nn;
    out[8] = s.V;
    out[9] = s.num_neurons;
  });
}

int gpuga_fitness_compute(gpuga_fitness* h, int P, const float* genes, float* fitness)
{
  return gpuga_fitness_compute_terms(h, P, genes, fitness, nullptr);
}

int gpuga_fitness_compute_terms(gpuga_fitness* h, int P, const float* genes, float* fitness,
                                double* terms)
{
  return guarded([&] {
    need(h, "handle");
    need(genes, "genes");
    need(fitness, "fitness");
    h->f->compute(P, genes, fitness, terms);
  });
}

int gpuga_fitness_compute_plain(gpuga_fitness* h, int P, const float* genes, float* fitness,
                                double* terms)
{
  return guarded([&] {
    need(h, "handle");
    need(genes, "genes");
    need(fitness, "fitness");
    h->f->compute_plain(P, genes, fitness, terms);
  });
}

int gpuga_fitness_compute_plain_noalloc(gpuga_fitness* h, int P, const float* genes,
                                        float* fitness, double* terms)
{
  return guarded([&] {
    need(h, "handle");
    need(genes, "genes");
    need(fitness, "fitness");
    h->f->compute_plain_noalloc(P, genes, fitness, terms);
  });
}

int gpuga_fitness_predict(gpuga_fitness* h, const char* dir, const float* elite)
{
  return guarded([&] {
    need(h, "handle");
    need(dir, "dir");
    need(elite, "elite");
    h->f->predict(dir, elite);
  });
}

int gpuga_fitness_find_force(gpuga_fitness* h, const float* params, float* f, float* virial,
                             float* pe)
{
  return guarded([&] {
    need(h, "handle");
    need(params, "params");
    h->f->find_force(params, f, virial, pe);
  });
}

int gpuga_fitness_neighbors(const gpuga_fitness* h, int* NN, int* start, int* NL)
{
  return guarded([&] {
    need(h, "handle");
    const NeighborList& nl = h->f->neighbors();
    if (NN) std::memcpy(NN, nl.NN.data(), sizeof(int) * nl.NN.size());
    if (start) std::memcpy(start, nl.start.data(), sizeof(int) * nl.start.size());
    if (NL) std::memcpy(NL, nl.NL.data(), sizeof(int) * nl.NL.size());
  });
}

int gpuga_inspect_neighbors(const char* dir, int* sizes, int* NN, int* start, int* NL)
{
  return guarded([&] {
    need(dir, "input_dir");
    need(sizes, "sizes");
    const PotentialSpec s = read_potential(dir);
    const TrainSet ts = read_train(dir);
    const NeighborList nl = build_neighbors(ts, s.cutoff);
    sizes[0] = ts.N;
    sizes[1] = nl.total;
    sizes[2] = nl.max_nn;
    sizes[3] = ts.Nc;
    if (NN) std::memcpy(NN, nl.NN.data(), sizeof(int) * nl.NN.size());
    if (start) std::memcpy(start, nl.start.data(), sizeof(int) * nl.start.size());
    if (NL) std::memcpy(NL, nl.NL.data(), sizeof(int) * nl.NL.size());
  });
}

namespace {
void put_tables(int* sizes, float* bond, float* tri, float* rec, const std::vector<float4>& b,
                const std::vector<float4>& t, const std::vector<float4>& r)
{
  sizes[0] = (int)b.size();
  sizes[1] = (int)t.size();
  sizes[2] = (int)r.size();
  if (bond) std::memcpy(bond, b.data(), sizeof(float4) * b.size());
  if (tri) std::memcpy(tri, t.data(), sizeof(float4) * t.size());
  if (rec) std::memcpy(rec, r.data(), sizeof(float4) * r.size());
}
} // namespace

int gpuga_fitness_geometry(const gpuga_fitness* h, int* sizes, float* bond, float* tri, float* rec)
{
  return guarded([&] {
    need(h, "handle");
    need(sizes, "sizes");
    std::vector<float4> b, t, r;
    h->f->device_geometry(b, t, r);
    put_tables(sizes, bond, tri, rec, b, t, r);
  });
}

int gpuga_inspect_geometry(const char* dir, int* sizes, float* bond, float* tri, float* rec)
{
  return guarded([&] {
    need(dir, "input_dir");
    need(sizes, "sizes");
    const PotentialSpec s = read_potential(dir);
    const TrainSet ts = read_train(dir);
    const NeighborList nl = build_neighbors(ts, s.cutoff);
    const bool tersoff = s.type == POT_TERSOFF1 || s.type == POT_TERSOFF2;
    Geometry g = build_geometry(ts, nl, tersoff);
    std::vector<float4> b(std::max(nl.total, 1), make_float4(0, 0, 0, 1)), t(1, make_float4(0, 0, 0, 0)), r;
    std::memcpy(b.data(), g.bond.data(), sizeof(float4) * g.bond.size());
    if (tersoff) {
      if (g.tri_total > 0) {
        t.resize(g.tri.size());
        std::memcpy(t.data(), g.tri.data(), sizeof(float4) * g.tri.size());
      }
      // the population kernel's records: bonds shorter than the largest R2
      // (the whole list when none is that long)
      const float reach = tersoff_reach(s);
      NeighborList pn;
      if (reach < s.cutoff && nl.total > 0) {
        std::vector<float> d(nl.total);
        for (int k = 0; k < nl.total; ++k) d[k] = g.bond[k].w;
        pn = prune_neighbors(nl, d, reach);
      }
      const bool pruned = pn.total < nl.total && pn.NN.size() == nl.NN.size();
      const NeighborList& rl = pruned ? pn : nl;
      if (const int M = record_slots(rl.max_nn)) {
        const Geometry pg = pruned ? build_geometry(ts, rl, true) : Geometry();
        const std::vector<Vec4> rv =
          build_records(rl, pruned ? pg : g, bond_types(ts, rl, true), M);
        r.resize(rv.size());
        std::memcpy(r.data(), rv.data(), sizeof(float4) * rv.size());
      }
    }
    put_tables(sizes, bond, tri, rec, b, t, r);
  });
}

int gpuga_fitness_autotune(gpuga_fitness* h, int P, int* chunk_cost)
{
  return guarded([&] {
    need(h, "handle");
    const int c = h->f->autotune(P);
    if (chunk_cost) *chunk_cost = c;
  });
}

int gpuga_fitness_launch_shape(gpuga_fitness* h, int P, int* chunk_cost, int* queues,
                               int* blocks_per_cu, int* packed)
{
  return guarded([&] {
    need(h, "handle");
    if (P < 1) throw std::invalid_argument("population size must be positive");
    int c = 0, q = 0, b = 0, g = 0;
    h->f->plan_shape(P, c, q, b, g);
    if (chunk_cost) *chunk_cost = c;
    if (queues) *queues = q;
    if (blocks_per_cu) *blocks_per_cu = b;
    if (packed) *packed = g > 1 ? 1 : 0;
  });
}

int gpuga_fitness_set_timing(gpuga_fitness* h, int on)
{
  return guarded([&] {
    need(h, "handle");
    h->f->set_timing(on != 0);
  });
}

int gpuga_fitness_timing(gpuga_fitness* h, double* kernel_ms, long* launches, int reset)
{
  return guarded([&] {
    need(h, "handle");
    if (kernel_ms) *kernel_ms = h->f->kernel_ms();
    if (launches) *launches = h->f->kernel_launches();
    if (reset) h->f->reset_timers();
  });
}

int gpuga_fitness_trace(gpuga_fitness* h, int P, const float* genes, double* rows, int* n_items)
{
  return guarded([&] {
    need(h, "handle");
    need(genes, "genes");
    need(n_items, "n_items");
    const std::vector<double> r = h->f->trace(P, genes);
    const int n = (int)(r.size() / 12);
    if (rows) std::memcpy(rows, r.data(), sizeof(double) * 12 * std::min(n, *n_items));
    *n_items = n;
  });
}

int gpuga_fitness_set_wait_limit(gpuga_fitness* h, int spins)
{
  return guarded([&] {
    need(h, "handle");
    h->f->set_spin_limit(spins);
  });
}

int gpuga_fitness_device_bytes(const gpuga_fitness* h, double* bytes)
{
  return guarded([&] {
    need(h, "handle");
    need(bytes, "bytes");
    *bytes = (double)h->f->device_bytes();
  });
}

// ------------------------------------------------------------------- GA

int gpuga_ga_create(int population_size, int parent_number, int V, int maximum_generation,
                    float mutation_rate, unsigned seed, gpuga_ga** out)
{
  return guarded([&] {
    need(out, "out");
    GAParams p;
    p.population_size = population_size;
    p.parent_number = parent_number;
    p.maximum_generation = maximum_generation;
    p.mutation_rate = mutation_rate;
    p.seed = seed;
    std::vector<float> lo(V, 0.0f), hi(V, 1.0f);
    auto h = new gpuga_ga;
    try {
      h->g = std::make_unique<GA>(p, V, lo, hi);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

void gpuga_ga_destroy(gpuga_ga* g) { delete g; }

int gpuga_ga_get_population(const gpuga_ga* g, float* genes)
{
  return guarded([&] {
    need(g, "handle");
    need(genes, "genes");
    const auto& p = g->g->population();
    std::memcpy(genes, p.data(), sizeof(float) * p.size());
  });
}

int gpuga_ga_set_population(gpuga_ga* g, const float* genes)
{
  return guarded([&] {
    need(g, "handle");
    need(genes, "genes");
    auto& p = g->g->population();
    std::memcpy(p.data(), genes, sizeof(float) * p.size());
  });
}

int gpuga_ga_step(gpuga_ga* g, int generation, const float* fitness, float* sorted)
{
  return guarded([&] {
    need(g, "handle");
    need(fitness, "fitness");
    auto& f = g->g->fitness();
    std::memcpy(f.data(), fitness, sizeof(float) * f.size());
    g->g->sort_population();
    if (sorted) std::memcpy(sorted, f.data(), sizeof(float) * f.size());
    g->g->crossover();
    g->g->mutation(generation);
  });
}

int gpuga_ga_step_device(gpuga_ga* g, int generation, const float* fitness, float* sorted)
{
  return guarded([&] {
    need(g, "handle");
    need(fitness, "fitness");
    GA& ga = *g->g;
    if (!g->st) GPUGA_HIP(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking));
    if (!g->dev) g->dev = std::make_unique<DeviceGA>(ga, nullptr, 1, g->st);
    g->dev->upload(ga, generation, g->st);
    g->fit.upload(fitness, (size_t)ga.params().population_size, g->st);
    g->dev->stage(ga, generation, 1, g->st);
    g->dev->step(0, g->fit.get(), g->st);
    g->dev->download(ga, 1, g->st);
    if (sorted) std::memcpy(sorted, ga.fitness().data(), sizeof(float) * ga.fitness().size());
  });
}

int gpuga_fitness_run_ga(gpuga_fitness* h, gpuga_ga* g, int first_generation, int generations,
                         float* best, float* elite)
{
  return guarded([&] {
    need(h, "handle");
    need(g, "ga");
    if (generations < 0) fail("generations must be >= 0");
    h->f->run_ga(*g->g, first_generation, generations, best, elite);
  });
}

void gpuga_default_convert_options(gpuga_convert_options* o)
{
  if (!o) return;
  const ConvertOptions d;
  for (int k = 0; k < 3; ++k) o->cells[k] = d.cells[k];
  o->delta_e = d.delta_e;
  o->stress_in_kbar = d.stress_in_kbar ? 1 : 0;
}

namespace {
void write_text(const char* path, const std::string& s)
{
  FILE* f = std::fopen(path, "wb");
  if (!f) fail(strfmt("Failed to open %s!", path));
  const size_t n = std::fwrite(s.data(), 1, s.size(), f);
  if (std::fclose(f) != 0 || n != s.size()) fail(strfmt("cannot write %s", path));
}
} // namespace

int gpuga_construct_train_in(const char* list_text, const char* dir, const char* out_path,
                             const gpuga_convert_options* opt)
{
  return guarded([&] {
    need(list_text, "list_text");
    need(out_path, "out_path");
    ConvertOptions o;
    if (opt) {
      for (int k = 0; k < 3; ++k) o.cells[k] = opt->cells[k];
      o.delta_e = opt->delta_e;
      o.stress_in_kbar = opt->stress_in_kbar != 0;
    }
    write_text(out_path, construct_train_in(list_text, dir ? dir : "", o));
  });
}

int gpuga_expand_train_in(const char* in_path, const char* out_path, const int* cells3)
{
  return guarded([&] {
    need(in_path, "in_path");
    need(out_path, "out_path");
    const int two[3] = {2, 2, 2};
    std::string text;
    {
      FILE* f = std::fopen(in_path, "rb");
      if (!f) fail(strfmt("Failed to open %s!", in_path));
      char buf[1 << 16];
      size_t n;
      while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
      std::fclose(f);
    }
    write_text(out_path, expand_train_in(text, cells3 ? cells3 : two));
  });
}

int gpuga_read_ga_in(const char* dir, int* G, int* P, int* parents, float* rate, unsigned* seed)
{
  return guarded([&] {
    need(dir, "input_dir");
    const GAParams p = read_ga(dir);
    if (G) *G = p.maximum_generation;
    if (P) *P = p.population_size;
    if (parents) *parents = p.parent_number;
    if (rate) *rate = p.mutation_rate;
    if (seed) *seed = p.seed;
  });
}

int gpuga_run(const char* input_dir, const gpuga_options* opt, double* seconds_per_generation)
{
  return guarded([&] {
    need(input_dir, "input_dir");
    const std::string dir(input_dir);
    Fitness fit(dir, to_options(opt));
    GA ga(read_ga(dir), fit.number_of_variables(), fit.spec().pmin, fit.spec().pmax);
    const auto t0 = std::chrono::steady_clock::now();
    const int P = ga.params().population_size, G = ga.params().maximum_generation;
    if (fit.options().backend == 1 || P > kDeviceGaMaxP) {
      // per-individual launches (or a population beyond the device GA):
      // the host loop of GA::compute
      ga.run(dir, [&](int n, const float* genes, float* fitness) {
        fit.compute(n, genes, fitness);
      });
    } else {
      // GA::compute (src/ga.cu:73-92) with the population and the GA step on
      // the device, in runs of generations; the ga.out records of each run
      // are written when it returns (the same lines GA::output writes)
      const std::string path = dir + "/ga.out";
      FILE* fid = std::fopen(path.c_str(), "w");
      if (!fid) fail(strfmt("Failed to open %s!", path.c_str()));
      const int V = ga.V(), run = 100;
      std::vector<float> best(run), elite((size_t)run * V);
      try {
        for (int g = 0; g < G; g += run) {
          const int n = std::min(run, G - g);
          fit.run_ga(ga, g, n, best.data(), elite.data());
          for (int i = 0; i < n; ++i) {
            write_ga_line(fid, g + i, best[i], &elite[(size_t)i * V], fit.spec().pmin,
                          fit.spec().pmax);
            if ((g + i + 1) % 10 == 0)
              write_ga_line(stdout, g + i + 1, best[i], &elite[(size_t)i * V], fit.spec().pmin,
                            fit.spec().pmax);
          }
          std::fflush(fid);
        }
      } catch (...) {
        std::fclose(fid);
        throw;
      }
      std::fclose(fid);
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds_per_generation)
      *seconds_per_generation = std::chrono::duration<double>(t1 - t0).count() /
                                ga.params().maximum_generation;
    fit.predict(dir, ga.population().data());
  });
}

// -------------------------------------------------------------- serving

int gpuga_server_create(double budget, int max_active, const gpuga_options* opt,
                        gpuga_server** out)
{
  return guarded([&] {
    need(out, "out");
    auto h = new gpuga_server;
    h->s = std::make_unique<Server>(budget, max_active, to_options(opt));
    *out = h;
  });
}

void gpuga_server_destroy(gpuga_server* s) { delete s; }

int gpuga_server_submit(gpuga_server* s, const char* dir, int write_outputs, int* job_id)
{
  return guarded([&] {
    need(s, "handle");
    need(dir, "input_dir");
    const int id = s->s->submit(dir, write_outputs != 0);
    if (job_id) *job_id = id;
  });
}

int gpuga_server_drain(gpuga_server* s)
{
  return guarded([&] {
    need(s, "handle");
    s->s->drain();
  });
}

int gpuga_server_poll(gpuga_server* s, int* active, int* queued, int* finished)
{
  return guarded([&] {
    need(s, "handle");
    s->s->poll();
    if (active) *active = s->s->active();
    if (queued) *queued = s->s->queued();
    if (finished) *finished = s->s->finished();
  });
}

int gpuga_server_job(const gpuga_server* s, int id, int* state, int* generations,
                     float* best_fitness, float* elite)
{
  return guarded([&] {
    need(s, "handle");
    s->s->job_info(id, state, generations, best_fitness, elite);
  });
}

int gpuga_server_latencies(const gpuga_server* s, double* ms, int* n)
{
  return guarded([&] {
    need(s, "handle");
    need(n, "n");
    const std::vector<double>& v = s->s->latencies();
    const int m = std::min<int>(*n, (int)v.size());
    if (ms)
      for (int i = 0; i < m; ++i) ms[i] = v[i];
    *n = (int)v.size();
  });
}

int gpuga_server_admissions(const gpuga_server* s, double* ms, int* n)
{
  return guarded([&] {
    need(s, "handle");
    need(n, "n");
    const std::vector<double>& v = s->s->admissions();
    const int m = std::min<int>(*n, (int)v.size());
    if (ms)
      for (int i = 0; i < m; ++i) ms[i] = v[i];
    *n = (int)v.size();
  });
}

} // extern "C"
