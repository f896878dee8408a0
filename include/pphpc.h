/*
 * pphpc.h -- C ABI of the MI355X-native PPHPC per-iteration step.
 *
 * A drop-in for the path of the reference OpenCL GPU simulator `pp_gpu`
 * (opencl/pp/pp_gpu.c): same parameter file, same seed/iteration contract,
 * same per-iteration statistics file. Plain C types only; no exceptions cross
 * this boundary; every entry point returns 0 or a negative PP_* code with the
 * values of the reference's enum pp_error_codes (opencl/pp/pp_common.in.h:87-104).
 *
 * Threading: a pp_sim_t handle is used by one host thread at a time. Several
 * handles may share one GPU. The handle owns all device memory; the caller
 * owns every host array passed in.
 */
#ifndef PPHPC_H
#define PPHPC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: the values of enum pp_error_codes (pp_common.in.h:87-104). */
#define PP_SUCCESS 0
#define PP_UNKNOWN_ARGS (-1)
#define PP_INVALID_ARGS (-2)
#define PP_LIBRARY_ERROR (-3)
#define PP_UNABLE_TO_OPEN_PARAMS_FILE (-4)
#define PP_INVALID_PARAMS_FILE (-5)
#define PP_UNABLE_SAVE_STATS (-6)
#define PP_OUT_OF_RESOURCES (-8)
/* No counterpart in the reference, whose 16-bit energy field wraps silently
 * (pp_gpu.cl:166-173): an agent's energy after its action exceeded
 * PP_ENERGY_MAX. The simulation (or, in a batch, only that replica) stops at
 * that iteration, as the oracle does. */
#define PP_ENERGY_OVERFLOW (-9)

/* Largest agent energy (the 24-bit energy field of the packed agent). */
#define PP_ENERGY_MAX 16777215u

/* Agent type codes: SHEEP_ID / WOLF_ID of pp_common.in.h:48-52. */
#define PP_SHEEP 0
#define PP_WOLF 1

/* Default capacity: the reference's default maximum number of agents
 * (PPG_DEFAULT_MAX_AGENTS, pp_gpu.c:57). */
#define PP_DEFAULT_MAX_AGENTS 16777216u

/* Replaces PPParameters (pp_common.in.h:129-157): same 13 fields, same order,
 * same meaning. grid_xy = grid_x * grid_y. */
typedef struct pp_params {
	uint32_t init_sheep;
	uint32_t sheep_gain_from_food;
	uint32_t sheep_reproduce_threshold;
	uint32_t sheep_reproduce_prob;
	uint32_t init_wolves;
	uint32_t wolves_gain_from_food;
	uint32_t wolves_reproduce_threshold;
	uint32_t wolves_reproduce_prob;
	uint32_t grass_restart;
	uint32_t grid_x;
	uint32_t grid_y;
	uint32_t grid_xy;
	uint32_t iters;
} pp_params_t;

/* Replaces PPStatistics (pp_common.in.h:109-124): same 7 fields, same order,
 * widened from cl_uint to exact 64-bit counts and sums. */
typedef struct pp_stats64 {
	uint64_t sheep;
	uint64_t wolves;
	uint64_t grass;     /* cells with live grass (countdown 0) */
	uint64_t sheep_en;  /* sum of sheep energy */
	uint64_t wolves_en; /* sum of wolf energy */
	uint64_t grass_en;  /* sum of grass countdowns over all cells */
	uint64_t errors;
} pp_stats64_t;

typedef struct pp_sim pp_sim_t;

/* Replaces pp_load_params (pp_common.c:39-195). Reads "KEY = value" records;
 * each of the 12 keys exactly once. Returns PP_SUCCESS,
 * PP_UNABLE_TO_OPEN_PARAMS_FILE (missing file) or PP_INVALID_PARAMS_FILE
 * (unknown key, repeated key, missing key). NULL path = "config.txt"
 * (PP_DEFAULT_PARAMS_FILE, pp_common.in.h:55). */
int pp_load_params(const char *path, pp_params_t *out);

/* Replaces pp_stats_save (pp_common.c:205-240): writes nrows lines of
 * "%llu\t%llu\t%llu\t%f\t%f\t%f\n" (sheep, wolves, grass alive, mean sheep
 * energy, mean wolf energy, mean countdown); a mean of an empty population is
 * 0. Means are exact sums divided in double precision. NULL path =
 * "stats.txt" (pp_common.in.h:58). Returns PP_SUCCESS or PP_UNABLE_SAVE_STATS. */
int pp_stats_save(const char *path, const pp_stats64_t *rows, uint32_t nrows,
	const pp_params_t *p);

/* pp_stats_save byte for byte as the reference writes it (pp_common.c:216-
 * 224): counts and energy sums as its 32-bit cl_uint fields (sums above
 * 2^32 - 1 wrap, as there), counts printed with "%d", and each mean a
 * single-precision quotient printed with "%f". For files compared with
 * ones pp_gpu wrote; pp_stats_save keeps exact sums and double means. */
int pp_stats_save_f32(const char *path, const pp_stats64_t *rows,
	uint32_t nrows, const pp_params_t *p);

/* Message for a PP_* code. */
const char *pp_strerror(int code);

/* Replaces the set-up half of pp_gpu's main (pp_gpu.c:1765-1847): device
 * selection (-d, pp_gpu.c:326), RNG seed (-r, :330), agent capacity (-m,
 * :338; 0 = max(4 * initial agents, PP_DEFAULT_MAX_AGENTS)). Allocates every
 * device buffer. device < 0 = current device. */
int pp_sim_create(const pp_params_t *p, uint64_t seed, uint64_t capacity,
	int device, pp_sim_t **out);

/* Replaces init_cell / init_agent and the first statistics gathering of
 * ppg_simulate (pp_gpu.cl:368-439, pp_gpu.c:619-642, :671-776): (re)sets the
 * state to iteration 0 and writes its statistics (row 0) to *stats0 (may be
 * NULL). */
int pp_sim_init(pp_sim_t *s, pp_stats64_t *stats0);

/* One iteration of the loop body of ppg_simulate (pp_gpu.c:671-935): grass,
 * move, sort, cell index, actions, statistics. Writes the statistics of the
 * new state to *out (may be NULL). Returns PP_OUT_OF_RESOURCES when the agents
 * after births exceed the capacity (pp_gpu.c:847-852, :895-900), and
 * PP_ENERGY_OVERFLOW when an agent's energy after its action exceeds
 * PP_ENERGY_MAX (checked first). */
int pp_sim_step(pp_sim_t *s, pp_stats64_t *out);

/* iters iterations with no host synchronisation between them; rows 1..iters
 * of the statistics go to out[0..iters-1] (may be NULL). On
 * PP_OUT_OF_RESOURCES or PP_ENERGY_OVERFLOW the simulation stops at the
 * failing iteration (see pp_sim_error_step). */
int pp_sim_run(pp_sim_t *s, uint32_t iters, pp_stats64_t *out);

/* Replaces the ppg_simulate + pp_stats_save sequence of main
 * (pp_gpu.c:1857-1866): init + p->iters iterations; rows 0..iters written to
 * out (iters + 1 rows; may be NULL). */
int pp_sim_simulate(pp_sim_t *s, pp_stats64_t *out);

/* Completed iterations, live agents of the current state, the iteration at
 * which an error stopped the run (or UINT32_MAX), and the sum over completed
 * iterations of the live agents at the start of each (agent-steps). */
uint32_t pp_sim_step_count(const pp_sim_t *s);
uint64_t pp_sim_num_agents(const pp_sim_t *s);
uint32_t pp_sim_error_step(const pp_sim_t *s);
uint64_t pp_sim_agent_steps(const pp_sim_t *s);
uint64_t pp_sim_capacity(const pp_sim_t *s);

/* Exact event sums over the completed iterations since the last init:
 * [0] agents at the start of each iteration (agent-steps), [1] agents alive
 * after the move, [2] agents after actions and births, [3] occupied cells,
 * [4] grass cells eaten, [5] births. Used to state each kernel's algorithmic
 * bytes. */
int pp_sim_counters(const pp_sim_t *s, uint64_t out[6]);

/* State dump, the counterpart of PPG_DUMP (pp_gpu.c:438-555): agents in slot
 * order (cell = y * grid_x + x, energy, type) and the grass countdown of every
 * cell. Each pointer may be NULL; arrays hold pp_sim_num_agents() agents and
 * grid_xy cells. */
int pp_sim_get_state(pp_sim_t *s, uint32_t *cell, int32_t *energy,
	uint8_t *type, uint32_t *grass);

/* Debug dump of the current state, the counterpart of the reference's
 * PPG_DUMP build option (pp_gpu.c:36-46; ppg_dump at :438-555). Appends one
 * block per call to agents_path:
 *     "\nIter %u, agents=%llu\n", then per agent, in state order,
 *     "[%4llu] (%4u, %4u) type=%u energy=%d\n"      (index, x, y, type, energy)
 * and to cells_path:
 *     "\nIteration %u\n", then per cell, row-major,
 *     "(%u, %u) -> (%llu, %llu) %s [Grass: %u]\n"
 * with the first and last index of the cell's agents ("(-, -)" for a cell
 * without agents), "More than 1 agent present" when they differ, and the
 * grass countdown (0 = alive). The iteration is the number of iterations done
 * (0 = the initial state). dump_type bit 0x10 keeps only the cells that hold
 * agents, as pp_gpu.c:521-523 does, with a blank line where cells were left
 * out; the state holds no dead agents, so bit 0x01 changes nothing. Either
 * path may be NULL; append = 0 truncates the files first. */
int pp_sim_dump(pp_sim_t *s, const char *agents_path, const char *cells_path,
	uint32_t dump_type, int append);

/* The formatter of pp_sim_dump, for a state held in host memory (n agents,
 * p->grid_xy cells). */
int pp_dump_state(const char *agents_path, const char *cells_path, int append,
	uint32_t iter, uint32_t dump_type, const pp_params_t *p, uint64_t n,
	const uint32_t *cell, const int32_t *energy, const uint8_t *type,
	const uint32_t *grass);

/* Per-kernel timing, the counterpart of the reference's profiler summary and
 * --prof-agg export (pp_gpu.c:318-321, :1868-1881; pp_common.c:248-260).
 * When enabled, each kernel launch of init/step/run is bracketed by HIP events
 * on the simulation's stream and no graph is used. */
int pp_sim_set_profiling(pp_sim_t *s, int enable);
/* Number of kernels tracked, and for kernel i its name, launch count and
 * total / average duration in nanoseconds. A name can cover more than one
 * kernel: "bucket_scan" is the count tables' clear before a radix-sorted
 * tick, or the prefix tables and bucket records (k_bprefix, k_brec) before
 * the bucket sort of a large layout. */
int pp_sim_prof_count(const pp_sim_t *s);
int pp_sim_prof_get(pp_sim_t *s, int i, const char **name, uint64_t *launches,
	double *total_ns);
int pp_sim_prof_reset(pp_sim_t *s);
/* "name"\ttotal_ns\trelative lines, as pp_export_prof_agg_info
 * (pp_common.c:248-260). */
int pp_sim_prof_export(pp_sim_t *s, const char *path);

/* Radix passes of the agent sort for this grid (8-bit digits of the cell
 * index) and the work-group / tile shape in use. */
int pp_sim_info(const pp_sim_t *s, uint32_t *sort_passes, uint32_t *sort_tile,
	uint32_t *action_tile, uint32_t *grid_move);

/* Sort of the agents by cell, as pp_batch_set_sort. */
int pp_sim_set_sort(pp_sim_t *s, uint32_t mode);

/* Releases all device memory of the handle. NULL is accepted. */
void pp_sim_destroy(pp_sim_t *s);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under stream/graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): the run of pp_sim_create + pp_sim_simulate +
 * pp_sim_destroy with the same results (rows 0..p->iters, iters + 1 rows,
 * into the caller's host array `rows`), every device buffer taken from a
 * fixed 12 MiB scratch region of the library's device memory instead of
 * allocated. It calls no hipMalloc, hipFree, pinned-host or other allocating
 * routine, and captures no graph (kernels go to the null stream of `device`,
 * < 0 = current). Limits: capacity (0 = 4 x initial agents) at most 262,144
 * agents, grid_xy at most 262,144 cells, grass_restart at most 4096; larger
 * sizes return PP_INVALID_ARGS with a message on stderr. Each GPU holds its
 * own scratch until the process ends and every call on it reuses (and
 * re-initialises) that one region: calls that overlap on one GPU race.
 * Returns 0 or a PP_* code as pp_sim_simulate. */
int pp_sim_simulate_noalloc(const pp_params_t *p, uint64_t seed,
	uint64_t capacity, int device, pp_stats64_t *rows);

/* Blocks until all work of the handle's stream has finished. */
int pp_sim_sync(pp_sim_t *s);

/* ------------------------------------------------------------------------
 * Batches of replicas. No counterpart in the reference, which runs one
 * simulation per process with its parameters compiled into the kernels as
 * -D constants (pp_gpu.c:1522-1574); here they are per-replica table
 * entries, so replicas of different grids, parameters, seeds and iteration
 * counts advance together, one set of kernel launches per tick. A replica in
 * a batch produces exactly the statistics rows it produces alone
 * (pp_sim_simulate with the same parameters and seed). A pp_sim_t is a batch
 * of one.
 *
 * Admission is by device footprint: a batch allocates its buffers once
 * (pp_batch_footprint bytes); a replica needs a free slot, grid_xy
 * contiguous free cells and an agent reservation (the most agents it may
 * hold after births; above it the replica fails with PP_OUT_OF_RESOURCES,
 * alone; an agent above PP_ENERGY_MAX fails it alone with
 * PP_ENERGY_OVERFLOW). pp_batch_add returns PP_OUT_OF_RESOURCES when these are not free
 * now; slots, cells and reservations come back when replicas finish.
 * ---------------------------------------------------------------------- */

typedef struct pp_batch pp_batch_t;

/* Replica states. */
#define PP_REPLICA_RUNNING 1
#define PP_REPLICA_DONE 2
#define PP_REPLICA_FAILED 3

/* Device bytes a batch with these limits allocates. 0 for invalid limits. */
uint64_t pp_batch_footprint(uint64_t agent_capacity, uint64_t cell_capacity,
	uint32_t max_replicas, uint32_t max_grass_restart);

/* A batch holding up to agent_capacity agents (< 2^29) over all replicas,
 * cell_capacity cells, max_replicas replicas at once, each with
 * grass_restart <= max_grass_restart. device < 0 = current device. */
int pp_batch_create(uint64_t agent_capacity, uint64_t cell_capacity,
	uint32_t max_replicas, uint32_t max_grass_restart, int device,
	pp_batch_t **out);

/* Admits a replica for p->iters iterations with RNG seed `seed` and agent
 * reservation agent_reserve (0 = 4 x initial agents). Its row 0 is written
 * at admission; *id names it until pp_batch_release. */
int pp_batch_add(pp_batch_t *b, const pp_params_t *p, uint64_t seed,
	uint64_t agent_reserve, uint32_t *id);

/* Advances every running replica by up to `ticks` iterations (each stops at
 * its own count), collecting their statistics rows; returns early once no
 * replica runs. */
int pp_batch_run(pp_batch_t *b, uint32_t ticks);

/* State (PP_REPLICA_*), completed iterations, error code (0 unless failed)
 * and agent-steps of replica id. */
int pp_batch_replica(pp_batch_t *b, uint32_t id, int *state, uint32_t *steps,
	int *error, uint64_t *agent_steps);

/* Statistics rows first .. first + n - 1 of replica id (row 0 = initial
 * state), as pp_sim_simulate writes them. */
int pp_batch_rows(pp_batch_t *b, uint32_t id, uint32_t first, uint32_t n,
	pp_stats64_t *out);

/* Current state of a running replica, as pp_sim_get_state (cells local to
 * its grid); *n_agents receives its agent count. */
int pp_batch_get_state(pp_batch_t *b, uint32_t id, uint32_t *n_agents,
	uint32_t *cell, int32_t *energy, uint8_t *type, uint32_t *grass);

/* Forgets a finished (done or failed) replica and its rows. */
int pp_batch_release(pp_batch_t *b, uint32_t id);

/* Free agent reservation, longest free run of cells, free slots, and
 * running replicas. */
int pp_batch_resources(const pp_batch_t *b, uint64_t *agents_free,
	uint32_t *cells_free_run, uint32_t *slots_free, uint32_t *active);

/* Iterations the batch has run. */
uint32_t pp_batch_tick(const pp_batch_t *b);

/* Batch-wide event sums, as pp_sim_counters. */
int pp_batch_counters(const pp_batch_t *b, uint64_t out[6]);

/* Sort passes, sort tile, action tile and emit grid in use. */
int pp_batch_info(const pp_batch_t *b, uint32_t *sort_passes,
	uint32_t *sort_tile, uint32_t *action_tile, uint32_t *grid_move);

/* Sort of the agents by cell: mode 0 = the radix sort every iteration;
 * 8..11 = from the second iteration after a change of the replicas, row
 * buckets (blocks of 2^mode cells of a grid row); 1 (the default) = row
 * buckets of 2^11 cells for a single replica, and for a batch with at most
 * 16384 of them or at least 64 agents per bucket (its replicas' agents as
 * last collected), else the radix sort. Results do not depend on it. */
int pp_batch_set_sort(pp_batch_t *b, uint32_t mode);

/* Radix sort launch shape: keys per thread (8 or 16; tile = 256 x that) and
 * the widest radix digit (8 or 11 bits). Results do not depend on it. */
int pp_batch_set_shape(pp_batch_t *b, uint32_t keys_per_thread,
	uint32_t max_digit_bits);

/* Launch shape, the counterpart of the reference's work sizes
 * (ppg_worksizes_compute, pp_gpu.c:1001-1158): the action and emit kernels
 * launch tile_rounds (1, 2, 4 or 8; default 4) rounds of the work-groups the
 * GPU holds at once, the bucket sort bucket_eighths eighths (1..8; default 8)
 * of a round. Results do not depend on it. */
int pp_batch_set_launch(pp_batch_t *b, uint32_t tile_rounds,
	uint32_t bucket_eighths);
int pp_batch_get_launch(const pp_batch_t *b, uint32_t *tile_rounds,
	uint32_t *bucket_eighths);

/* Autotuner: runs trial_ticks iterations of the batch under each candidate
 * sort (bucket widths 2^8..2^11, radix shapes; the iterations count as
 * normal progress), keeps the fastest and reports it (bucket_shift 0 = radix
 * sort) with its seconds per tick. When the fastest is what the automatic
 * mode (pp_batch_set_sort 1) picks for the current replicas, the batch stays
 * in the automatic mode; otherwise the winner is pinned as if set with
 * pp_batch_set_sort / pp_batch_set_shape. Then, under that sort, each
 * launch shape (tile_rounds 1, 2, 4, 8; bucket_eighths 4, 8) is timed and
 * the fastest kept (pp_batch_get_launch reports it). A trial during which
 * replicas finish is timed again. */
int pp_batch_tune(pp_batch_t *b, uint32_t trial_ticks, uint32_t *bucket_shift,
	uint32_t *keys_per_thread, uint32_t *max_digit_bits, double *sec_per_tick);

/* Per-kernel timing of the batch, as pp_sim_set_profiling / prof_get. */
int pp_batch_set_profiling(pp_batch_t *b, int enable);
int pp_batch_prof_count(const pp_batch_t *b);
int pp_batch_prof_get(pp_batch_t *b, int i, const char **name,
	uint64_t *launches, double *total_ns);
int pp_batch_prof_reset(pp_batch_t *b);

int pp_batch_sync(pp_batch_t *b);

/* Releases all device memory of the batch. NULL is accepted. */
void pp_batch_destroy(pp_batch_t *b);

#ifdef __cplusplus
}
#endif

#endif
