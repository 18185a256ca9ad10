// Serving loop (host).
#include "serve.h"

#include <algorithm>
#include <cstring>
#include <thread>

namespace gpuga {

Server::Server(double budget_bytes, int max_active, const FitnessOptions& opt)
  : budget_(budget_bytes), max_active_(std::max(1, max_active)), opt_(opt)
{
  GPUGA_HIP(hipGetDevice(&device_));
}

Server::~Server()
{
  for (auto& j : jobs_)
    if (j->task.valid()) j->task.wait();
  for (auto& j : jobs_)
    if (j->gaout) std::fclose(j->gaout);
}

int Server::submit(const std::string& dir, bool write_outputs)
{
  auto j = std::make_unique<Job>();
  j->t_submit = std::chrono::steady_clock::now();
  j->id = (int)jobs_.size();
  j->dir = dir;
  j->write = write_outputs;
  // the job's files are read (and its device bytes estimated) on a worker
  Job* jp = j.get();
  const FitnessOptions opt = opt_;
  const int dev = device_;
  j->task = std::async(std::launch::async, [jp, opt, dev] {
    GPUGA_HIP(hipSetDevice(dev));
    jp->spec = read_potential(jp->dir);
    jp->ts = read_train(jp->dir);
    jp->gp = read_ga(jp->dir);
    jp->est_bytes = Fitness::estimate_bytes(jp->spec, jp->ts, jp->gp.population_size, opt);
  });
  queue_.push_back(j->id);
  jobs_.push_back(std::move(j));
  return jobs_.back()->id;
}

bool Server::task_done(Job& j) const
{
  return j.task.wait_for(std::chrono::seconds(0)) == std::future_status::ready;
}

int Server::active() const
{
  int n = 0;
  for (auto& j : jobs_) n += j->state == 1;
  return n;
}

int Server::finished() const
{
  int n = 0;
  for (auto& j : jobs_) n += j->state >= 2;
  return n;
}

void Server::admit(Job& j)
{
  // the job counts as active (slot and memory) from here; its set-up runs
  // on a worker and the job starts generations once that has finished
  j.state = 1;
  j.ready = false;
  used_ += (double)j.est_bytes;
  Job* jp = &j;
  const FitnessOptions opt = opt_;
  const int dev = device_;
  j.task = std::async(std::launch::async, [jp, opt, dev] {
    GPUGA_HIP(hipSetDevice(dev));
    Job& w = *jp;
    FitnessOptions o = opt;
    w.fit = std::make_unique<Fitness>(w.spec, std::move(w.ts), o);
    // deterministic work-item size: the same chunking alone or co-scheduled
    if (o.chunk_cost <= 0) w.fit->options().chunk_cost = w.fit->heuristic_chunk_cost(w.gp.population_size);
    // plan and graph now: generation latencies then measure generations only
    w.fit->prepare(w.gp.population_size);
    w.ga = std::make_unique<GA>(w.gp, w.fit->number_of_variables(), w.fit->spec().pmin,
                                w.fit->spec().pmax);
    if (w.write) {
      const std::string path = w.dir + "/ga.out";
      w.gaout = std::fopen(path.c_str(), "w");
      if (!w.gaout) fail(strfmt("Failed to open %s!", path.c_str()));
    }
  });
}

void Server::launch(Job& j, int share)
{
  j.t_launch = std::chrono::steady_clock::now();
  j.fit->compute_async(j.gp.population_size, j.ga->population().data(), share);
  j.inflight = true;
}

void Server::complete(Job& j)
{
  // GA::compute body for one generation (src/ga.cu:82-88)
  j.fit->compute_wait(j.ga->fitness().data(), nullptr);
  j.inflight = false;
  const auto now = std::chrono::steady_clock::now();
  lat_ms_.push_back(std::chrono::duration<double, std::milli>(now - j.t_launch).count());
  j.ga->sort_population();
  if (j.gaout) {
    j.ga->write_line(j.gaout, j.gen);
    std::fflush(j.gaout);
  }
  j.ga->crossover();
  j.ga->mutation(j.gen);
  ++j.gen;
  if (j.gen >= j.gp.maximum_generation) retire(j, true);
}

void Server::retire(Job& j, bool ok, const std::string& err)
{
  if (ok) {
    j.best = j.ga->fitness()[0];
    const int V = j.ga->V();
    j.elite.resize(V);
    for (int m = 0; m < V; ++m) {
      const float a = j.fit->spec().pmin[m];
      const float b = j.fit->spec().pmax[m] - a;
      const float bx = b * j.ga->population()[m];
      j.elite[m] = a + bx;
    }
    if (j.write) {
      try {
        j.fit->predict(j.dir, j.ga->population().data());
      } catch (const std::exception& e) {
        ok = false;
        j.error = e.what();
      }
    }
  } else {
    j.error = err;
  }
  if (j.gaout) {
    std::fclose(j.gaout);
    j.gaout = nullptr;
  }
  if (j.state == 1) used_ -= (double)j.est_bytes;
  j.state = ok ? 2 : 3;
  j.fit.reset();
  j.ga.reset();
}

void Server::poll()
{
  // 1. consume finished generations
  for (auto& jp : jobs_) {
    Job& j = *jp;
    if (j.state != 1 || !j.ready || !j.inflight) continue;
    const hipError_t q = hipStreamQuery(j.fit->stream());
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) {
      retire(j, false, hipGetErrorString(q));
      continue;
    }
    try {
      complete(j);
    } catch (const std::exception& e) {
      retire(j, false, e.what());
    }
  }
  // 2. jobs whose set-up has finished become ready (admission time: from
  // submission to ready, waiting for the budget included)
  for (auto& jp : jobs_) {
    Job& j = *jp;
    if (j.state != 1 || j.ready || !task_done(j)) continue;
    try {
      j.task.get();
      j.ready = true;
      adm_ms_.push_back(
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - j.t_submit)
          .count());
    } catch (const std::exception& e) {
      retire(j, false, e.what());
    }
  }
  // 3. admit queued jobs in order while a slot and the memory budget allow
  while (!queue_.empty()) {
    Job& j = *jobs_[queue_.front()];
    if (!j.parsed) {
      if (!task_done(j)) break; // still reading its files
      try {
        j.task.get();
        j.parsed = true;
      } catch (const std::exception& e) {
        queue_.pop_front();
        j.state = 3;
        j.error = e.what();
        continue;
      }
    }
    if (active() >= max_active_) break;
    if (active() > 0 && used_ + (double)j.est_bytes > budget_) break;
    queue_.pop_front();
    admit(j);
  }
  // 4. start the next generation of every idle ready job, with serve_share on
  // its share of the GPU (the jobs running generations split the device);
  // results do not depend on the share
  // (at most 4: launches from more streams than the process's hardware
  // queues, 4 by default, do not run at once)
  int running = 0;
  for (auto& jp : jobs_) running += jp->state == 1 && jp->ready;
  const int share = opt_.serve_share ? std::max(1, std::min(running, 4)) : 1;
  for (auto& jp : jobs_) {
    Job& j = *jp;
    if (j.state == 1 && j.ready && !j.inflight) {
      try {
        launch(j, share);
      } catch (const std::exception& e) {
        retire(j, false, e.what());
      }
    }
  }
}

void Server::drain()
{
  while (finished() < (int)jobs_.size()) {
    const size_t before = lat_ms_.size();
    poll();
    if (lat_ms_.size() == before) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

void Server::job_info(int id, int* state, int* gens, float* best, float* elite) const
{
  if (id < 0 || id >= (int)jobs_.size()) fail(strfmt("no job %d", id));
  const Job& j = *jobs_[id];
  if (j.state == 3 && !j.error.empty() && state) {
    *state = 3;
  } else if (state) {
    *state = j.state;
  }
  if (gens) *gens = j.gen;
  if (best) *best = j.best;
  if (elite && !j.elite.empty()) std::memcpy(elite, j.elite.data(), sizeof(float) * j.elite.size());
  if (j.state == 3) fail(strfmt("job %d failed: %s", id, j.error.c_str()));
}

} // namespace gpuga
