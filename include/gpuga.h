/*
 * gpuga — C ABI of the MI355X fitness-evaluation library.
 *
 * Every entry point returns 0 on success and non-zero on failure, with the
 * message available from gpuga_last_error() (per calling thread).  Arrays are
 * plain host pointers; no entry point keeps a pointer it was given.  One
 * handle must not be used by two host threads at the same time.
 *
 * Each declaration names the reference interface it replaces
 * (brucefan1983/GPUGA, file:line).
 */
#ifndef GPUGA_H
#define GPUGA_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpuga_fitness gpuga_fitness;
typedef struct gpuga_ga gpuga_ga;
typedef struct gpuga_server gpuga_server;

typedef struct gpuga_options {
  int backend;    /* 0: one population launch per generation (default);
                     1: per-individual launches, as src/fitness.cu:318-332 */
  int chunk_cost; /* work-item size of the population launch; 0 = autotune
                     the launch shape (this and the fields below left at 0
                     or, for pack_lanes, at 1) */
  int use_graph;  /* 1: replay the population launch from a HIP graph */
  int nn2b_bf16;  /* 1: NN2B hidden layer on BF16 MFMA; 0: FP32 */
  int nn2b_grid;  /* NN2B tabulation points (default 1024) */
  int verbose;
  int blocks_per_cu; /* resident workgroups per CU of the population launch;
                        0 = as many as the occupancy calculator allows (or
                        one fewer, when the autotuner finds that faster) */
  int pack_lanes;    /* 1: a last lane block of <= 16 individuals runs several
                        atoms per wave (three-body potentials; the autotuner
                        may decide against it); 0: one atom */
  int queues;        /* work queues of the population launch; 0 = four per
                        XCD (or 16 / 64, as the autotuner finds faster) */
  int ga_batch;      /* generations per staged batch of gpuga_fitness_run_ga's
                        random decisions; 0 = up to 32 */
  int tail_split;    /* configurations smaller than one work item are cut into
                        pieces of 1/tail_split of an item (0, 1: not cut) */
  int serve_share;   /* server: 1 = each running job's launches on 1/n of the
                        GPU for n running jobs (up to 4); 0 = the whole GPU
                        (default: measured faster) */
} gpuga_options;

/* Thread-local text of the last failure (src/error.cu:21-25 prints it). */
const char* gpuga_last_error(void);
void gpuga_default_options(gpuga_options* opt);

/* Fitness::Fitness(char* input_dir), src/fitness.cu:30-38: reads
 * potential.in and train.in, builds the neighbour list, keeps the training
 * set resident on the current HIP device. */
int gpuga_fitness_create(const char* input_dir, const gpuga_options* opt, gpuga_fitness** out);

/* The same from arrays in train.in order: Na[Nc], type[N], box9[Nc][9] (file
 * rows a, b, c), r[3][N], pe_ref[Nc], virial_ref[6][Nc], force_ref[3][N];
 * pmin/pmax[V] as potential.in (ignored for NN2B: [-3, 3]), weights as
 * potential.in lines 3-5 (src/fitness.cu:267-298). */
int gpuga_fitness_create_arrays(int potential_type, int num_neurons, float cutoff,
                                float weight_force, float weight_energy, float weight_virial,
                                const float* pmin, const float* pmax, int Nc, int Nc_force,
                                const int* Na, const int* type, const float* box9, const float* r,
                                const float* pe_ref, const float* virial_ref,
                                const float* force_ref, const gpuga_options* opt,
                                gpuga_fitness** out);

void gpuga_fitness_destroy(gpuga_fitness* h);

/* Fitness::number_of_variables and parameters_min/max, src/fitness.cuh:35-38 */
int gpuga_fitness_num_variables(const gpuga_fitness* h);
int gpuga_fitness_bounds(const gpuga_fitness* h, float* pmin, float* pmax);

/* Sizes: out[0..9] = potential_type, Nc, Nc_force, N, N_force, max_Na,
 * bonds, max neighbours, V, num_neurons (src/fitness.cuh:55-63). */
int gpuga_fitness_info(const gpuga_fitness* h, int* out10);

/* Fitness::compute(int P, float* genes, float* fitness), src/fitness.cu:316-333.
 * genes: [P][V] in [0,1], row-major; fitness: [P].  Blocking. */
int gpuga_fitness_compute(gpuga_fitness* h, int P, const float* genes, float* fitness);

/* As above, also returning terms[P][3] = relative RMS error of energy,
 * virial, force (get_fitness_energy/stress/force, src/fitness.cu:437-555). */
int gpuga_fitness_compute_terms(gpuga_fitness* h, int P, const float* genes, float* fitness,
                                double* terms);

/* The per-individual-launch form of Fitness::compute (src/fitness.cu:318-332):
 * one find_force, one reduction and one host synchronisation per individual. */
int gpuga_fitness_compute_plain(gpuga_fitness* h, int P, const float* genes, float* fitness,
                                double* terms);

/* gpuga_fitness_compute_plain with every device scratch array (the
 * per-bond and per-atom arrays of find_force, the per-configuration errors)
 * in a fixed workspace of 64 MiB per GPU, a file-scope device array set up
 * with the handle: the call allocates and frees no device memory.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops).  Fails
 * when the training set needs more than 64 MiB of workspace (5 floats per
 * bond for minimal Tersoff, 10 floats per atom, 3 doubles per
 * configuration).  All calls on one GPU share its workspace: calls that
 * overlap on one GPU (two threads, two handles) race. */
int gpuga_fitness_compute_plain_noalloc(gpuga_fitness* h, int P, const float* genes,
                                        float* fitness, double* terms);

/* Fitness::predict(char* dir, float* elite), src/fitness.cu:347-390: writes
 * force.out, energy.out, virial.out into dir. */
int gpuga_fitness_predict(gpuga_fitness* h, const char* dir, const float* elite_genes);

/* Potential::update_potential + Potential::find_force, src/potential.cuh:26-40,
 * for physical parameters params[V]: f[3][N], virial[6][N], pe[N]
 * (any output may be NULL). */
int gpuga_fitness_find_force(gpuga_fitness* h, const float* params, float* f, float* virial,
                             float* pe);

/* Neighbour list of Neighbor::compute, src/neighbor.cu:30-83, in CSR form:
 * NN[N], start[N], NL[bonds] (ascending j per atom). */
int gpuga_fitness_neighbors(const gpuga_fitness* h, int* NN, int* start, int* NL);

/* Host-only set-up check (no HIP device used): reads potential.in and
 * train.in from dir and builds the neighbour list of Neighbor::compute
 * (src/neighbor.cu:30-66).  sizes[0..3] = N, bonds, max neighbours, Nc; call
 * once with NULL arrays for the sizes, then with NN[N], start[N], NL[bonds]. */
int gpuga_inspect_neighbors(const char* input_dir, int* sizes4, int* NN, int* start, int* NL);

/* ----------------------------------------------------------- training-set tools
 * The reference's offline train.in tools, host only, with their compile-time
 * constants as options and byte-identical output.  list_text is what the
 * converter reads on standard input (Nc_force Nc_energy, then the DFT file
 * names, relative to dir); the result is written to out_path. */
typedef struct gpuga_convert_options {
  int cells[3];       /* energy/virial cell replicated NA x NB x NC (default 4 4 3) */
  float delta_e;      /* energy shift per atom in eV (default 1.31684) */
  int stress_in_kbar; /* 1: stress in kbar (VASP) converted to a virial in eV */
} gpuga_convert_options;
void gpuga_default_convert_options(gpuga_convert_options* opt);
/* tools/construct_train.in/construct_train_in.cpp:62-222 (main, get_Nc,
 * get_files, get_Na, get_force, get_energy_and_virial) */
int gpuga_construct_train_in(const char* list_text, const char* dir, const char* out_path,
                             const gpuga_convert_options* opt);
/* tools/construct_train.in/expand222.cpp:41-151: every configuration of
 * in_path replicated cells[0] x cells[1] x cells[2] (NULL: 2 x 2 x 2). */
int gpuga_expand_train_in(const char* in_path, const char* out_path, const int* cells3);

/* Inspection of the parameter-independent geometry tables the population
 * kernels read (the reference recomputes this geometry in every kernel,
 * src/minimal_tersoff.cu:150-153,162-167, src/mic.cuh:20-32).  Arrays of
 * float4: bond[sizes3[0]] (r_ij, |r_ij|), tri[sizes3[1]] (cos theta_jik and
 * its gradient), rec[sizes3[2]] (Tersoff atom records).  Call once with NULL
 * arrays for the sizes.  gpuga_fitness_geometry copies the tables built on
 * the device; gpuga_inspect_geometry builds them on the host (no HIP device
 * used) from potential.in and train.in. */
int gpuga_fitness_geometry(const gpuga_fitness* h, int* sizes3, float* bond, float* tri, float* rec);
int gpuga_inspect_geometry(const char* input_dir, int* sizes3, float* bond, float* tri, float* rec);

/* Launch shape chosen by timing candidates for population P (work-item
 * size, then the queue count, resident workgroups per CU and lane packing
 * that the options leave at 0 / on); returns the work-item size.  No
 * counterpart in the reference (its launch shape is fixed). */
int gpuga_fitness_autotune(gpuga_fitness* h, int P, int* chunk_cost);
/* The shape the population launch for P runs with (tuned first when the
 * options ask for it): work-item cost, queues, resident workgroups per CU
 * and 1 if the last lane block is packed.  Any pointer may be NULL.  No
 * counterpart in the reference. */
int gpuga_fitness_launch_shape(gpuga_fitness* h, int P, int* chunk_cost, int* queues,
                               int* blocks_per_cu, int* packed);

/* HIP-event timing of the population launches (or of each individual's
 * launches on the per-individual backend) on the library's stream. */
int gpuga_fitness_set_timing(gpuga_fitness* h, int on);
int gpuga_fitness_timing(gpuga_fitness* h, double* kernel_ms, long* launches, int reset);

/* One population generation with its work-item timeline recorded (no
 * reference counterpart; a profiling aid).  rows[n_items][12] = configuration,
 * lane block, chunk, first atom, end atom, phase, wave slot, start ns, end ns
 * (when the wave's next fetch returned), for a wave's first item the wave's
 * entry ns (else NaN), then the ns at which the item's compute began (its
 * parameters set up, or its dependency wait over) and ended (before the
 * reductions); times relative to the first start.  Call with rows = NULL to
 * get *n_items. */
int gpuga_fitness_trace(gpuga_fitness* h, int P, const float* genes, double* rows, int* n_items);

/* Polls of the population kernel's dependency wait (a force-assembly item
 * waiting for the partial forces of its configuration) before the launch
 * reports a timeout: compute then fails, the launch's results are discarded
 * and the handle stays usable.  Default about a second; < 0 reports a
 * timeout at once (no reference counterpart: a test hook for the failure
 * path, the reference has no dependency waits). */
int gpuga_fitness_set_wait_limit(gpuga_fitness* h, int spins);

/* Device bytes held by the handle (training set, tables, plans). */
int gpuga_fitness_device_bytes(const gpuga_fitness* h, double* bytes);

/* ---------------------------------------------------------------- GA
 * class GA, src/ga.cu:27-212, with the random stream of std::mt19937 and
 * libstdc++'s distributions (explicit algorithms). */
int gpuga_ga_create(int population_size, int parent_number, int V, int maximum_generation,
                    float mutation_rate, unsigned seed, gpuga_ga** out);
void gpuga_ga_destroy(gpuga_ga* g);
/* genes [P][V] (copy out / in) */
int gpuga_ga_get_population(const gpuga_ga* g, float* genes);
int gpuga_ga_set_population(gpuga_ga* g, const float* genes);
/* After fitness[P] of the current population: sort_population, crossover,
 * mutation(generation) (src/ga.cu:83-88).  sorted_fitness[P] may be NULL. */
int gpuga_ga_step(gpuga_ga* g, int generation, const float* fitness, float* sorted_fitness);

/* The same step computed by the device GA kernel, with identical results
 * (no reference counterpart; it is what gpuga_fitness_run_ga runs each
 * generation, exposed so that it can be checked against gpuga_ga_step).
 * Populations of at most 4096. */
int gpuga_ga_step_device(gpuga_ga* g, int generation, const float* fitness, float* sorted_fitness);

/* GA::compute, src/ga.cu:73-92: `generations` generations (Fitness::compute,
 * sort_population, crossover, mutation) starting at generation
 * first_generation, with the population and the GA step on the device and
 * no host round trip in between; the GA handle holds the state before and
 * after (bit-identical to gpuga_fitness_compute + gpuga_ga_step per
 * generation).  best[generations] and elite[generations][V] (genes in [0, 1])
 * receive the ga.out record of each generation (src/ga.cu:138-159); either
 * may be NULL.  Populations of at most 4096. */
int gpuga_fitness_run_ga(gpuga_fitness* h, gpuga_ga* g, int first_generation, int generations,
                         float* best, float* elite);

/* ga.in reader of GA::read_parameters (src/ga.cu:214-316); host only.  Also
 * accepts an optional `seed` keyword (default 12345678, the reference's
 * deterministic seed, src/ga.cu:47-48). */
int gpuga_read_ga_in(const char* input_dir, int* maximum_generation, int* population_size,
                     int* parent_number, float* mutation_rate, unsigned* seed);

/* GPUGA::GPUGA(char* dir), src/gpuga.cu:24-28: reads ga.in, potential.in,
 * train.in from dir, runs the GA, writes ga.out, force.out, energy.out,
 * virial.out.  seconds_per_generation may be NULL. */
int gpuga_run(const char* input_dir, const gpuga_options* opt, double* seconds_per_generation);

/* ----------------------------------------------------------- serving
 * Concurrent fitting jobs (no reference counterpart: the reference runs the
 * directories of input.txt one after another, src/main.cu:28-50).  Jobs are
 * admitted while their device bytes fit the budget; each admitted job keeps
 * its training set and population resident and advances one generation at a
 * time on its own stream, so generations of different jobs overlap. */
int gpuga_server_create(double memory_budget_bytes, int max_active, const gpuga_options* opt,
                        gpuga_server** out);
void gpuga_server_destroy(gpuga_server* s);
/* Queue a job directory (ga.in, potential.in, train.in); returns its id. */
int gpuga_server_submit(gpuga_server* s, const char* input_dir, int write_outputs, int* job_id);
/* Run the scheduler until every submitted job has finished. */
int gpuga_server_drain(gpuga_server* s);
/* Advance the scheduler by at most one generation per active job. */
int gpuga_server_poll(gpuga_server* s, int* active, int* queued, int* finished);
/* Job results: state (0 queued, 1 active, 2 done, 3 failed), generations done,
 * best fitness, elite physical parameters[V] (may be NULL). */
int gpuga_server_job(const gpuga_server* s, int job_id, int* state, int* generations,
                     float* best_fitness, float* elite);
/* Per-generation latencies (ms) of all jobs in completion order; n in/out. */
int gpuga_server_latencies(const gpuga_server* s, double* ms, int* n);
/* Per admitted job, submission to ready (parsing, set-up, work plan, graph;
 * queueing behind the memory budget included), ms, in admission order. */
int gpuga_server_admissions(const gpuga_server* s, double* ms, int* n);

#ifdef __cplusplus
}
#endif
#endif
