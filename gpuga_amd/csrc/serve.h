// Serving loop: concurrent GA fitting jobs with resident training sets.
//
// The reference runs the job directories of input.txt one after another
// (src/main.cu:28-50).  Here jobs are queued, admitted while their device
// bytes fit a budget (and a slot is free), and each admitted job advances one
// generation at a time on its own HIP stream, so generations of different
// jobs overlap on the device.  A finished job releases its memory and the
// next queued job is admitted (continuous batching).  Optionally
// (serve_share) the jobs running generations split the device: each launch
// runs on 1/n of its grid for n running jobs (up to 4).  By default every
// launch takes the whole GPU: launches from up to four streams overlap at
// their edges, and that measured faster than splitting.  Reading a job's files
// and setting it up (training set on the device, work plan, graph) run on
// worker threads, so admissions never hold up the generations of running
// jobs.  The work-item size of each job is a deterministic function of its
// training set, so a job's results are the same whether it runs alone or
// with others.
#pragma once

#include <chrono>
#include <cstdio>
#include <deque>
#include <future>
#include <memory>
#include <string>
#include <vector>

#include "fitness.h"
#include "ga.h"

namespace gpuga {

class Server {
public:
  Server(double budget_bytes, int max_active, const FitnessOptions& opt);
  ~Server();
  int submit(const std::string& dir, bool write_outputs);
  void poll();
  void drain();
  int active() const;
  int queued() const { return (int)queue_.size(); }
  int finished() const;
  void job_info(int id, int* state, int* gens, float* best, float* elite) const;
  const std::vector<double>& latencies() const { return lat_ms_; }
  // per admitted job: submission to ready (parse, set-up, plan, graph), ms
  const std::vector<double>& admissions() const { return adm_ms_; }

private:
  struct Job {
    int id = 0;
    std::string dir;
    bool write = false;
    int state = 0; // 0 queued, 1 active (set-up or running), 2 done, 3 failed
    bool parsed = false; // files read and device bytes estimated
    bool ready = false;  // set-up finished: generations can run
    std::string error;
    PotentialSpec spec;
    TrainSet ts;
    GAParams gp;
    size_t est_bytes = 0;
    std::unique_ptr<Fitness> fit;
    std::unique_ptr<GA> ga;
    FILE* gaout = nullptr;
    int gen = 0;
    bool inflight = false;
    std::chrono::steady_clock::time_point t_launch, t_submit;
    float best = 0.0f;
    std::vector<float> elite;
    // the worker task (reading, then set-up); last member, so that it is
    // waited for before the others are destroyed
    std::future<void> task;
  };
  void admit(Job& j);
  bool task_done(Job& j) const;
  void launch(Job& j, int share);
  void complete(Job& j);
  void retire(Job& j, bool ok, const std::string& err = "");

  double budget_;
  int max_active_;
  int device_ = 0;
  FitnessOptions opt_;
  std::vector<std::unique_ptr<Job>> jobs_;
  std::deque<int> queue_;
  double used_ = 0.0;
  std::vector<double> lat_ms_, adm_ms_;
};

} // namespace gpuga
