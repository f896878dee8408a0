/*
 * pp_oracle.c -- single-threaded CPU oracle for the PPHPC per-iteration step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation. Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark use it.
 *
 * Model semantics restated here (reference files, relative to the reference
 * repository root):
 *   - init of cells:   netlogo/pphpc.nlogo:27-35, opencl/pp/pp_gpu.cl:368-398,
 *                      opencl/pp/pp_cpu.cl:382-406
 *   - init of agents:  netlogo/pphpc.nlogo:38-53, opencl/pp/pp_gpu.cl:406-439
 *   - move:            opencl/pp/pp_gpu.cl:744-821 (direction table 1:+x 2:-x
 *                      3:+y 4:-y, torus wrap), netlogo/pphpc.nlogo:66-70
 *   - grass regrowth:  opencl/pp/pp_gpu.cl:448-476, netlogo/pphpc.nlogo:151-159
 *   - actions:         java/src/org/laseeb/pphpc/Cell.java:236-259 (shuffle,
 *                      act alive agents in order), Sheep.java:66-83,
 *                      Wolf.java:66-92 (first alive sheep of the shuffled list),
 *                      AbstractAgent.java:139-167 (reproduction),
 *                      opencl/pp/pp_cpu.cl:690-845 (newborns never eaten)
 *   - statistics:      java/src/org/laseeb/pphpc/Cell.java:175-231,
 *                      netlogo/pphpc.nlogo:100-114
 *   - parameter I/O:   opencl/pp/pp_common.c:39-195 and :205-240
 *
 * Deterministic choices shared with the GPU implementation (the reference
 * draws these from stateful per-work-item generators, so no reference stream
 * can be matched bit for bit):
 *   - RNG: Philox4x32-10, key = (seed low 32 bits, seed high 32 bits),
 *     counter = (agent id or cell, step, stream, block). Bounded draws use
 *     the multiply-high map floor(x * n / 2^32).
 *   - Agent ids at step s are the slot order of the state after step s-1.
 *   - Agents of a cell are listed in ascending id order (a stable sort by
 *     cell), then permuted by a Fisher-Yates shuffle keyed by (cell, step),
 *     and act in that order.
 *   - The state after a step lists cells in ascending order; within a cell,
 *     surviving agents in acting order, each followed by its newborn.
 *
 * Parity status of each function is in the repository design notes; every
 * function here is checked by the CPU test suite against known answers,
 * rocRAND, brute force or hand-worked cases.
 */
#define _POSIX_C_SOURCE 200809L
#include "pp_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Streams of the counter-based RNG (third counter word). */
#define PPO_STREAM_INIT_CELL 0u
#define PPO_STREAM_INIT_AGENT 1u
#define PPO_STREAM_MOVE 2u
#define PPO_STREAM_SHUFFLE 3u

#define PPO_SHEEP 0
#define PPO_WOLF 1

/* Largest energy an agent may hold: 2^24 - 1, the energy field of the GPU's
 * packed agent value. An agent whose energy after its action (eating, then
 * reproduction) exceeds it ends the simulation with PPO_ENERGY_OVERFLOW at
 * that iteration, before the capacity check. The reference stores energy in
 * a 16-bit field that wraps silently (pp_gpu.cl:166-173, :251-258); here the
 * limit is explicit and its failure is a distinct code. */
#define PPO_ENERGY_MAX 0x00FFFFFF
#define PPO_ENERGY_OVERFLOW (-9)

/* ------------------------------------------------------------------ RNG */

static inline uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t *hi) {
	uint64_t p = (uint64_t)a * (uint64_t)b;
	*hi = (uint32_t)(p >> 32);
	return (uint32_t)p;
}

void ppo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
	uint32_t out[4]) {
	/* Round function and key schedule of Philox4x32 as published in
	 * "Parallel random numbers: as easy as 1, 2, 3" (Salmon, Moraes, Dror,
	 * Shaw, SC'11), 10 rounds. */
	const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
	const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	uint32_t k0 = key[0], k1 = key[1];
	for (int r = 0; r < 10; r++) {
		uint32_t hi0, hi1;
		uint32_t lo0 = mulhilo32(M0, c0, &hi0);
		uint32_t lo1 = mulhilo32(M1, c2, &hi1);
		uint32_t n0 = hi1 ^ c1 ^ k0;
		uint32_t n2 = hi0 ^ c3 ^ k1;
		c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
		k0 += W0; k1 += W1;
	}
	out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform integer in [0, n) from a 32-bit word: floor(x * n / 2^32). */
static inline uint32_t bounded(uint32_t x, uint32_t n) {
	return (uint32_t)(((uint64_t)x * (uint64_t)n) >> 32);
}

static inline void draw(const uint32_t key[2], uint32_t a, uint32_t step,
	uint32_t stream, uint32_t block, uint32_t out[4]) {
	uint32_t ctr[4] = {a, step, stream, block};
	ppo_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------- params I/O */

int ppo_load_params(const char *path, ppo_params *p) {
	/* Restates pp_load_params (pp_common.c:39-195): "%s = %d" records, the
	 * twelve keys each exactly once, unknown key or repeat -> -5, missing
	 * file -> -4; grid_xy derived (pp_common.c:177). */
	static const char *names[12] = {
		"INIT_SHEEP", "SHEEP_GAIN_FROM_FOOD", "SHEEP_REPRODUCE_THRESHOLD",
		"SHEEP_REPRODUCE_PROB", "INIT_WOLVES", "WOLVES_GAIN_FROM_FOOD",
		"WOLVES_REPRODUCE_THRESHOLD", "WOLVES_REPRODUCE_PROB",
		"GRASS_RESTART", "GRID_X", "GRID_Y", "ITERS"};
	uint32_t vals[12] = {0};
	unsigned seen = 0;
	char name[100];
	int value;
	FILE *fp = fopen(path, "r");
	if (!fp) return -4;
	while (fscanf(fp, "%99s = %d", name, &value) != EOF) {
		int k;
		for (k = 0; k < 12; k++)
			if (strcmp(name, names[k]) == 0) break;
		if (k == 12 || (seen & (1u << k))) { fclose(fp); return -5; }
		seen |= 1u << k;
		vals[k] = (uint32_t)value;
	}
	fclose(fp);
	if (seen != 0xfffu) return -5;
	p->init_sheep = vals[0];
	p->sheep_gain_from_food = vals[1];
	p->sheep_reproduce_threshold = vals[2];
	p->sheep_reproduce_prob = vals[3];
	p->init_wolves = vals[4];
	p->wolves_gain_from_food = vals[5];
	p->wolves_reproduce_threshold = vals[6];
	p->wolves_reproduce_prob = vals[7];
	p->grass_restart = vals[8];
	p->grid_x = vals[9];
	p->grid_y = vals[10];
	p->grid_xy = vals[9] * vals[10];
	p->iters = vals[11];
	return 0;
}

int ppo_stats_save(const char *path, const ppo_stats *rows, uint32_t nrows,
	uint32_t grid_xy) {
	/* Restates pp_stats_save (pp_common.c:217-224): sheep, wolves, grass
	 * alive, mean sheep energy, mean wolf energy, mean countdown; a mean is 0
	 * for an empty population. Means come from exact 64-bit sums. */
	FILE *fp = fopen(path, "w");
	if (!fp) return -6;
	for (uint32_t i = 0; i < nrows; i++) {
		const ppo_stats *r = &rows[i];
		fprintf(fp, "%llu\t%llu\t%llu\t%f\t%f\t%f\n",
			(unsigned long long)r->sheep, (unsigned long long)r->wolves,
			(unsigned long long)r->grass,
			r->sheep ? (double)r->sheep_en / (double)r->sheep : 0.0,
			r->wolves ? (double)r->wolves_en / (double)r->wolves : 0.0,
			(double)r->grass_en / (double)grid_xy);
	}
	fclose(fp);
	return 0;
}

/* ------------------------------------------------------------ simulation */

struct ppo_sim {
	ppo_params p;
	uint32_t key[2];
	uint64_t cap;
	uint32_t step;
	uint64_t n;
	uint64_t agent_steps;
	uint64_t events[5];
	int fault;
	/* state */
	uint32_t *cell;
	int32_t *energy;
	uint8_t *type;
	uint32_t *grass;
	/* scratch */
	uint32_t *mcell;   /* cell after move */
	int32_t *menergy;  /* energy after move */
	uint8_t *mdice;    /* reproduction draw, [0,100) */
	uint8_t *malive;   /* survived the move */
	uint32_t *count;   /* per-cell counts, C + 1 */
	uint32_t *sorted;  /* ids sorted by cell, stable */
	uint8_t *killed;
	int32_t *child_e;  /* -1: no child */
	uint32_t *ocell;
	int32_t *oenergy;
	uint8_t *otype;
};

static void *xcalloc(size_t n, size_t sz) {
	void *p = calloc(n ? n : 1, sz);
	if (!p) { fprintf(stderr, "pp_oracle: out of memory\n"); abort(); }
	return p;
}

ppo_sim *ppo_create(const ppo_params *p, uint64_t seed, uint64_t capacity) {
	uint64_t n0 = (uint64_t)p->init_sheep + p->init_wolves;
	if (p->grid_x == 0 || p->grid_y == 0 || p->grass_restart == 0)
		return NULL;
	if (capacity < n0) return NULL;
	ppo_sim *s = (ppo_sim *)xcalloc(1, sizeof(ppo_sim));
	s->p = *p;
	s->p.grid_xy = p->grid_x * p->grid_y;
	s->key[0] = (uint32_t)seed;
	s->key[1] = (uint32_t)(seed >> 32);
	s->cap = capacity;
	uint64_t C = s->p.grid_xy;
	s->cell = xcalloc(capacity, 4);
	s->energy = xcalloc(capacity, 4);
	s->type = xcalloc(capacity, 1);
	s->grass = xcalloc(C, 4);
	s->mcell = xcalloc(capacity, 4);
	s->menergy = xcalloc(capacity, 4);
	s->mdice = xcalloc(capacity, 1);
	s->malive = xcalloc(capacity, 1);
	s->count = xcalloc(C + 1, 4);
	s->sorted = xcalloc(capacity, 4);
	s->killed = xcalloc(capacity, 1);
	s->child_e = xcalloc(capacity, 4);
	s->ocell = xcalloc(2 * capacity, 4);
	s->oenergy = xcalloc(2 * capacity, 4);
	s->otype = xcalloc(2 * capacity, 1);

	/* Cells: alive with probability 1/2, else countdown uniform in
	 * [1, GRASS_RESTART] (pp_gpu.cl:382-390; pphpc.nlogo:27-35). */
	for (uint64_t c = 0; c < C; c++) {
		uint32_t r[4];
		draw(s->key, (uint32_t)c, 0, PPO_STREAM_INIT_CELL, 0, r);
		if (bounded(r[0], 2) == 1)
			s->grass[c] = 0;
		else
			s->grass[c] = 1 + bounded(r[1], p->grass_restart);
	}
	/* Agents: sheep first, then wolves; uniform cell; energy uniform in
	 * [1, 2 * gain] (pp_gpu.cl:417-434; pphpc.nlogo:38-53). */
	for (uint64_t i = 0; i < n0; i++) {
		uint32_t r[4];
		int t = i < p->init_sheep ? PPO_SHEEP : PPO_WOLF;
		uint32_t gain = t == PPO_SHEEP ? p->sheep_gain_from_food
			: p->wolves_gain_from_food;
		draw(s->key, (uint32_t)i, 0, PPO_STREAM_INIT_AGENT, 0, r);
		uint32_t x = bounded(r[0], p->grid_x);
		uint32_t y = bounded(r[1], p->grid_y);
		s->cell[i] = y * p->grid_x + x;
		s->energy[i] = (int32_t)(1 + bounded(r[2], 2 * gain));
		s->type[i] = (uint8_t)t;
	}
	s->n = n0;
	return s;
}

void ppo_destroy(ppo_sim *s) {
	if (!s) return;
	free(s->cell); free(s->energy); free(s->type); free(s->grass);
	free(s->mcell); free(s->menergy); free(s->mdice); free(s->malive);
	free(s->count); free(s->sorted); free(s->killed); free(s->child_e);
	free(s->ocell); free(s->oenergy); free(s->otype);
	free(s);
}

void ppo_set_fault(ppo_sim *s, int mode) { s->fault = mode; }

void ppo_stats_get(const ppo_sim *s, ppo_stats *out) {
	/* Post-step statistics: all agents present (newborns included), grass
	 * alive = countdown 0, countdown sum over all cells (Cell.java:175-231,
	 * pphpc.nlogo:100-114). */
	memset(out, 0, sizeof(*out));
	for (uint64_t i = 0; i < s->n; i++) {
		if (s->type[i] == PPO_SHEEP) {
			out->sheep++;
			out->sheep_en += (uint64_t)s->energy[i];
		} else {
			out->wolves++;
			out->wolves_en += (uint64_t)s->energy[i];
		}
	}
	for (uint64_t c = 0; c < s->p.grid_xy; c++) {
		if (s->grass[c] == 0) out->grass++;
		out->grass_en += s->grass[c];
	}
}

static uint32_t move_cell(const ppo_params *p, uint32_t cell, uint32_t dir) {
	/* Direction table of the reference GPU move (pp_gpu.cl:780-803):
	 * 0 stay, 1 +x, 2 -x, 3 +y, 4 -y, all with torus wrap. */
	uint32_t x = cell % p->grid_x, y = cell / p->grid_x;
	switch (dir) {
	case 1: x = (x + 1 == p->grid_x) ? 0 : x + 1; break;
	case 2: x = (x == 0) ? p->grid_x - 1 : x - 1; break;
	case 3: y = (y + 1 == p->grid_y) ? 0 : y + 1; break;
	case 4: y = (y == 0) ? p->grid_y - 1 : y - 1; break;
	default: break;
	}
	return y * p->grid_x + x;
}

int ppo_step(ppo_sim *s) {
	const ppo_params *p = &s->p;
	const uint32_t st = s->step;
	const uint64_t C = p->grid_xy;
	const uint64_t n = s->n;

	/* 1. Move: direction and reproduction draw from one Philox block keyed
	 * by (id, step); energy -1; death at energy < 1 (pphpc.nlogo:66-70). */
	for (uint64_t i = 0; i < n; i++) {
		uint32_t r[4];
		draw(s->key, (uint32_t)i, st, PPO_STREAM_MOVE, 0, r);
		uint32_t dir = bounded(r[0], 5);
		if (s->fault == 2 && (dir == 1 || dir == 2)) dir = 3 - dir;
		s->mdice[i] = (uint8_t)bounded(r[1], 100);
		s->menergy[i] = s->energy[i] - 1;
		s->malive[i] = s->menergy[i] >= 1;
		s->mcell[i] = move_cell(p, s->cell[i], dir);
	}

	/* 2. Grass regrowth: countdown -1 where positive (pp_gpu.cl:465). */
	for (uint64_t c = 0; c < C; c++)
		if (s->grass[c] > 0) s->grass[c]--;

	/* 3. Stable counting sort of surviving agents by cell. */
	memset(s->count, 0, (C + 1) * sizeof(uint32_t));
	for (uint64_t i = 0; i < n; i++)
		if (s->malive[i]) s->count[s->mcell[i] + 1]++;
	for (uint64_t c = 0; c < C; c++) s->count[c + 1] += s->count[c];
	uint64_t live = s->count[C];
	for (uint64_t i = 0; i < n; i++)
		if (s->malive[i]) s->sorted[s->count[s->mcell[i]]++] = (uint32_t)i;

	/* 4. Per-cell actions in shuffled order, emitting the next state. */
	uint64_t ev[5] = {n - live, 0, 0, 0, 0};
	uint64_t out = 0;
	int over = 0;
	uint64_t q0 = 0;
	while (q0 < live) {
		uint32_t c = s->mcell[s->sorted[q0]];
		uint64_t q1 = q0 + 1;
		while (q1 < live && s->mcell[s->sorted[q1]] == c) q1++;
		uint32_t *A = &s->sorted[q0];
		uint32_t k = (uint32_t)(q1 - q0);

		/* Fisher-Yates (Durstenfeld) from the last position down, the draw
		 * for position i being word m % 4 of block m / 4, m = k-1-i, of the
		 * stream keyed by (cell, step) (shuffle of Cell.java:246 and
		 * pp_cpu.cl:236-259). */
		if (s->fault != 3) {
			uint32_t r[4];
			for (uint32_t i = k - 1, m = 0; i > 0; i--, m++) {
				if ((m & 3u) == 0)
					draw(s->key, c, st, PPO_STREAM_SHUFFLE, m >> 2, r);
				uint32_t j = bounded(r[m & 3u], i + 1);
				uint32_t t = A[i]; A[i] = A[j]; A[j] = t;
			}
		}

		for (uint32_t a = 0; a < k; a++) {
			s->killed[A[a]] = 0;
			s->child_e[A[a]] = -1;
		}
		for (uint32_t a = 0; a < k; a++) {
			uint32_t id = A[a];
			if (s->killed[id]) continue;
			int32_t e = s->menergy[id];
			uint32_t thr, prob;
			if (s->type[id] == PPO_SHEEP) {
				/* Sheep eat live grass (Sheep.java:66-83). */
				if (s->grass[c] == 0) {
					s->grass[c] = p->grass_restart;
					e += (int32_t)p->sheep_gain_from_food
						* (s->fault == 4 ? 2 : 1);
					ev[1]++;
				}
				thr = p->sheep_reproduce_threshold;
				prob = p->sheep_reproduce_prob;
			} else {
				/* Wolves eat the first live sheep of the shuffled list,
				 * acted or not; newborns are not in the list
				 * (Wolf.java:66-92, pp_cpu.cl:718-763). */
				for (uint32_t b = 0; b < k; b++) {
					uint32_t o = A[b];
					if (s->type[o] == PPO_SHEEP && !s->killed[o]) {
						s->killed[o] = 1;
						e += (int32_t)p->wolves_gain_from_food
							+ (s->fault == 1 ? 1 : 0);
						ev[2]++;
						break;
					}
				}
				thr = p->wolves_reproduce_threshold;
				prob = p->wolves_reproduce_prob;
			}
			/* Reproduction: child gets floor(e/2), parent keeps the rest
			 * (AbstractAgent.java:139-167, pphpc.nlogo:133-142). */
			if (e > (int32_t)thr && s->mdice[id] < prob) {
				int32_t ce = e / 2;
				e -= ce;
				s->child_e[id] = ce;
				ev[3 + s->type[id]]++;
			}
			over |= e > PPO_ENERGY_MAX;
			s->menergy[id] = e;
		}
		for (uint32_t a = 0; a < k; a++) {
			uint32_t id = A[a];
			if (!s->killed[id]) {
				s->ocell[out] = c;
				s->oenergy[out] = s->menergy[id];
				s->otype[out] = s->type[id];
				out++;
			}
			if (s->child_e[id] >= 0) {
				s->ocell[out] = c;
				s->oenergy[out] = s->child_e[id];
				s->otype[out] = s->type[id];
				out++;
			}
		}
		q0 = q1;
	}

	if (over) return PPO_ENERGY_OVERFLOW;
	if (out > s->cap) {
		/* PP_OUT_OF_RESOURCES (pp_gpu.c:847-852). The grass pass already
		 * ran; the error ends the simulation, as in the reference. */
		return -8;
	}
	memcpy(s->cell, s->ocell, out * sizeof(uint32_t));
	memcpy(s->energy, s->oenergy, out * sizeof(int32_t));
	memcpy(s->type, s->otype, out);
	s->agent_steps += n;
	s->n = out;
	s->step++;
	memcpy(s->events, ev, sizeof(ev));
	return 0;
}

int ppo_state_set(ppo_sim *s, uint64_t n, const uint32_t *cell,
	const int32_t *energy, const uint8_t *type, const uint32_t *grass) {
	if (n > s->cap) return -8;
	memcpy(s->cell, cell, n * sizeof(uint32_t));
	memcpy(s->energy, energy, n * sizeof(int32_t));
	memcpy(s->type, type, n);
	memcpy(s->grass, grass, (size_t)s->p.grid_xy * sizeof(uint32_t));
	s->n = n;
	return 0;
}

void ppo_last_events(const ppo_sim *s, uint64_t out[5]) {
	memcpy(out, s->events, sizeof(s->events));
}

int ppo_run(ppo_sim *s, uint32_t iters, ppo_stats *out) {
	for (uint32_t t = 0; t < iters; t++) {
		int rc = ppo_step(s);
		if (rc) return rc;
		if (out) ppo_stats_get(s, &out[t]);
	}
	return 0;
}

uint32_t ppo_step_count(const ppo_sim *s) { return s->step; }
uint64_t ppo_num_agents(const ppo_sim *s) { return s->n; }
uint64_t ppo_agent_steps(const ppo_sim *s) { return s->agent_steps; }

void ppo_state_get(const ppo_sim *s, uint32_t *cell, int32_t *energy,
	uint8_t *type, uint32_t *grass) {
	if (cell) memcpy(cell, s->cell, s->n * sizeof(uint32_t));
	if (energy) memcpy(energy, s->energy, s->n * sizeof(int32_t));
	if (type) memcpy(type, s->type, s->n);
	if (grass) memcpy(grass, s->grass, (size_t)s->p.grid_xy * sizeof(uint32_t));
}

/* Text dump of the state (the format of pp_sim_dump, after the reference's
 * PPG_DUMP output, opencl/pp/pp_gpu.c:438-555): one block per call appended
 * to each file; bit 0x10 of `type` lists only the cells holding agents. */
int ppo_dump(const ppo_sim *s, const char *agents_path, const char *cells_path,
	uint32_t type, int append) {
	const char *mode = append ? "a" : "w";
	const uint32_t gx = s->p.grid_x, C = s->p.grid_xy;
	if (agents_path) {
		FILE *f = fopen(agents_path, mode);
		if (!f) return -6;
		fprintf(f, "\nIter %u, agents=%llu\n", s->step,
			(unsigned long long)s->n);
		for (uint64_t i = 0; i < s->n; i++)
			fprintf(f, "[%4llu] (%4u, %4u) type=%u energy=%d\n",
				(unsigned long long)i, s->cell[i] % gx, s->cell[i] / gx,
				(unsigned)s->type[i], s->energy[i]);
		fclose(f);
	}
	if (cells_path) {
		/* first and last slot of each cell (after init the state is in
		 * creation order, after a step in cell order) */
		uint64_t *lo = xcalloc(C, sizeof(uint64_t)), *hi = xcalloc(C, sizeof(uint64_t));
		for (uint32_t c = 0; c < C; c++) lo[c] = UINT64_MAX;
		for (uint64_t i = 0; i < s->n; i++) {
			const uint32_t c = s->cell[i];
			if (lo[c] == UINT64_MAX) lo[c] = i;
			hi[c] = i;
		}
		FILE *f = fopen(cells_path, mode);
		if (!f) {
			free(lo);
			free(hi);
			return -6;
		}
		fprintf(f, "\nIteration %u\n", s->step);
		int gap = 0;
		for (uint32_t c = 0; c < C; c++) {
			const int empty = lo[c] == UINT64_MAX;
			if (empty && (type & 0x10)) {
				gap = 1;
				continue;
			}
			if (gap) fputc('\n', f);
			gap = 0;
			if (empty)
				fprintf(f, "(%u, %u) -> (-, -)  [Grass: %u]\n", c % gx, c / gx,
					s->grass[c]);
			else
				fprintf(f, "(%u, %u) -> (%llu, %llu) %s [Grass: %u]\n", c % gx,
					c / gx, (unsigned long long)lo[c], (unsigned long long)hi[c],
					hi[c] > lo[c] ? "More than 1 agent present" : "", s->grass[c]);
		}
		fclose(f);
		free(lo);
		free(hi);
	}
	return 0;
}

/* ------------------------------------------------------------------- CLI */
#ifdef PPO_MAIN
#include <getopt.h>

static double now_s(void) {
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
	/* Same flags as the reference executable (pp_gpu.c:310-345): -p params
	 * (default config.txt), -s stats (default stats.txt), -r seed (default
	 * 0), -m capacity; plus -n to override ITERS, -q to skip the file, and
	 * -D PREFIX [-T TYPE] to dump every state to PREFIX_agents.txt and
	 * PREFIX_cells.txt. */
	const char *params = "config.txt", *stats = "stats.txt";
	uint64_t seed = 0, cap = 0;
	long iters_override = -1;
	const char *dump = NULL;
	uint32_t dump_type = 0x01;
	int quiet = 0, opt;
	while ((opt = getopt(argc, argv, "p:s:r:m:n:qD:T:")) != -1) {
		switch (opt) {
		case 'p': params = optarg; break;
		case 's': stats = optarg; break;
		case 'r': seed = strtoull(optarg, NULL, 0); break;
		case 'm': cap = strtoull(optarg, NULL, 0); break;
		case 'n': iters_override = strtol(optarg, NULL, 0); break;
		case 'q': quiet = 1; break;
		case 'D': dump = optarg; break;
		case 'T': dump_type = (uint32_t)strtoul(optarg, NULL, 0); break;
		default: return -1;
		}
	}
	ppo_params p;
	int rc = ppo_load_params(params, &p);
	if (rc) { fprintf(stderr, "Error: parameters file (%d)\n", rc); return rc; }
	if (iters_override >= 0) p.iters = (uint32_t)iters_override;
	uint64_t n0 = (uint64_t)p.init_sheep + p.init_wolves;
	if (!cap) cap = n0 * 4 > 16777216 ? n0 * 4 : 16777216;
	ppo_sim *s = ppo_create(&p, seed, cap);
	if (!s) { fprintf(stderr, "Error: invalid arguments\n"); return -2; }
	ppo_stats *rows = xcalloc((size_t)p.iters + 1, sizeof(ppo_stats));
	ppo_stats_get(s, &rows[0]);
	double t0 = now_s();
	if (dump) {
		char *fa = xcalloc(strlen(dump) + 16, 1), *fc = xcalloc(strlen(dump) + 16, 1);
		sprintf(fa, "%s_agents.txt", dump);
		sprintf(fc, "%s_cells.txt", dump);
		rc = ppo_dump(s, fa, fc, dump_type, 0);
		for (uint32_t t = 0; !rc && t < p.iters; t++) {
			rc = ppo_run(s, 1, rows + 1 + t);
			if (!rc) rc = ppo_dump(s, fa, fc, dump_type, 1);
		}
		free(fa);
		free(fc);
	} else {
		rc = ppo_run(s, p.iters, rows + 1);
	}
	double t1 = now_s();
	if (rc) fprintf(stderr, "Error: %s at step %u\n",
		rc == PPO_ENERGY_OVERFLOW ? "energy above its limit" : "out of resources",
		ppo_step_count(s));
	printf("agent_steps=%llu seconds=%.6f agent_steps_per_s=%.6e\n",
		(unsigned long long)ppo_agent_steps(s), t1 - t0,
		(double)ppo_agent_steps(s) / (t1 - t0));
	if (!quiet && !rc) {
		int w = ppo_stats_save(stats, rows, p.iters + 1, p.grid_xy);
		if (w) rc = w;
	}
	free(rows);
	ppo_destroy(s);
	return rc;
}
#endif
