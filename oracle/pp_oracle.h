/*
 * pp_oracle.h -- single-threaded CPU oracle for the PPHPC per-iteration step.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the HIP implementation is
 * compared against: only the tests, the smoke check and the CPU-baseline leg of
 * the benchmark may load it. The product library never links or calls it.
 *
 * It restates the model semantics of the reference (see pp_oracle.c for the
 * file:line each function follows) with a deterministic ordering and a
 * counter-based RNG keyed by (seed, step, agent id or cell), so that every
 * per-iteration statistic of the GPU implementation can be checked
 * bit-exactly.
 */
#ifndef PP_ORACLE_H
#define PP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 13 fields, same order as PPParameters (pp_common.in.h:129-157). */
typedef struct ppo_params {
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
} ppo_params;

/* PPStatistics (pp_common.in.h:109-124) widened to 64-bit exact sums. */
typedef struct ppo_stats {
	uint64_t sheep;
	uint64_t wolves;
	uint64_t grass;
	uint64_t sheep_en;
	uint64_t wolves_en;
	uint64_t grass_en;
	uint64_t errors;
} ppo_stats;

typedef struct ppo_sim ppo_sim;

/* Philox4x32-10 (Salmon et al., SC'11): one block for counter ctr, key key. */
void ppo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
	uint32_t out[4]);

/* Parameter file parser restating pp_load_params (pp_common.c:39-195).
 * Returns 0, -4 (cannot open) or -5 (invalid file). */
int ppo_load_params(const char *path, ppo_params *p);

/* Stats writer restating pp_stats_save (pp_common.c:205-240) with exact
 * 64-bit sums and double means. Returns 0 or -6. */
int ppo_stats_save(const char *path, const ppo_stats *rows, uint32_t nrows,
	uint32_t grid_xy);

/* Simulation handle. capacity = maximum live agents (after births). */
ppo_sim *ppo_create(const ppo_params *p, uint64_t seed, uint64_t capacity);
void ppo_destroy(ppo_sim *s);

/* Stats of the current state (row 0 right after ppo_create). */
void ppo_stats_get(const ppo_sim *s, ppo_stats *out);

/* Advance one iteration. Returns 0; -9 (PP_ENERGY_OVERFLOW) when an agent's
 * energy after its action exceeds 2^24 - 1; or -8 (PP_OUT_OF_RESOURCES) when
 * the agents after births exceed capacity. The simulation cannot continue
 * after either (the agent arrays keep the previous state, the grass does
 * not). */
int ppo_step(ppo_sim *s);

/* Run iters iterations from the current state, writing rows 1..iters of the
 * stats into out[0..iters-1]. Returns 0, -9 or -8 (and stops at that step). */
int ppo_run(ppo_sim *s, uint32_t iters, ppo_stats *out);

/* Number of completed steps and live agents. */
uint32_t ppo_step_count(const ppo_sim *s);
uint64_t ppo_num_agents(const ppo_sim *s);

/* Copy the state: agents in slot (id) order and the grass countdown per
 * cell. Any pointer may be NULL. */
void ppo_state_get(const ppo_sim *s, uint32_t *cell, int32_t *energy,
	uint8_t *type, uint32_t *grass);

/* Replace the state: n agents (cell, energy, type) in slot order and the
 * grass countdown of every cell. Returns 0 or -8 if n exceeds capacity. */
int ppo_state_set(ppo_sim *s, uint64_t n, const uint32_t *cell,
	const int32_t *energy, const uint8_t *type, const uint32_t *grass);

/* Event counts of the last completed step: [0] deaths by starvation,
 * [1] grass eaten, [2] sheep eaten, [3] sheep births, [4] wolf births. */
void ppo_last_events(const ppo_sim *s, uint64_t out[5]);

/* Appends the text dump of the state (format of pp_sim_dump) to the files;
 * either may be NULL; append = 0 truncates them. 0 or a negative code. */
int ppo_dump(const ppo_sim *s, const char *agents_path, const char *cells_path,
	uint32_t type, int append);

/* Sum over all steps of the live agents at the start of each step. */
uint64_t ppo_agent_steps(const ppo_sim *s);

/* Fault injection for testing the tests: 0 = none. A non-zero mode plants a
 * deliberate mistake in the step so that parity tests can show they catch
 * it (1: wolf gain off by one, 2: move direction 1 and 2 swapped,
 * 3: shuffle skipped, 4: sheep gain doubled). */
void ppo_set_fault(ppo_sim *s, int mode);

#ifdef __cplusplus
}
#endif

#endif
