/*
 * pp_omp.c -- multi-threaded CPU restatement of the PPHPC step (OpenMP).
 *
 * BASELINE AND TEST INFRASTRUCTURE ONLY: the CPU baseline that the benchmark
 * times beside the GPU, and a second CPU implementation the tests compare
 * with the single-threaded oracle. Nothing in the product links it.
 *
 * Parallel layout after the reference's CPU simulator (opencl/pp/pp_cpu.c:
 * 230-334, opencl/pp/pp_cpu.cl:477-866): the grid is split into horizontal
 * bands of rows, one per thread, and each thread runs the grass and the
 * agent actions of the cells of its own band. The step semantics and the
 * random streams are those of the oracle (pp_oracle.c, counter-based
 * Philox4x32-10 keyed by seed; streams by agent id or cell), so the result
 * does not depend on the thread count and equals the oracle's bit for bit:
 *
 *   1. move      threads take contiguous runs of agent ids; each surviving
 *                agent is counted in the band of its new cell
 *   2. partition the survivors are scattered to their bands, keeping id
 *                order inside each band (runs in thread order)
 *   3. band work each thread, for its band: grass countdown, stable counting
 *                sort of its agents by cell, then per cell in ascending order
 *                the oracle's shuffle and actions, writing the band's part of
 *                the next state to a buffer of its own
 *   4. gather    the bands' parts are concatenated in band order (= cell
 *                order, the oracle's order) and the statistics summed
 *
 *   pp_omp -p FILE [-s FILE] [-r SEED] [-m CAP] [-n ITERS] [-t THREADS] [-q]
 *          [-b FILE]
 *
 * prints "agent_steps=N seconds=S agent_steps_per_s=R threads=T" like the
 * oracle CLI, and writes the statistics file unless -q. -b FILE also writes
 * the exact rows computed (rows 0 .. last completed iteration, each seven
 * little-endian uint64 in ppo_stats order), also after a failure.
 */
#define _POSIX_C_SOURCE 200809L
#include "pp_oracle.h"

#include <getopt.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Streams of the counter-based RNG (third counter word), as the oracle's. */
enum { S_INIT_CELL = 0, S_INIT_AGENT = 1, S_MOVE = 2, S_SHUFFLE = 3 };

static inline uint32_t bounded(uint32_t x, uint32_t n) {
	return (uint32_t)(((uint64_t)x * n) >> 32);
}

static inline void philox(const uint32_t key[2], uint32_t a, uint32_t step,
	uint32_t stream, uint32_t block, uint32_t out[4]) {
	const uint32_t ctr[4] = {a, step, stream, block};
	ppo_philox4x32_10(ctr, key, out);
}

static void *xalloc(size_t n) {
	void *p = malloc(n ? n : 1);
	if (!p) {
		fprintf(stderr, "pp_omp: out of memory\n");
		exit(3);
	}
	return p;
}

typedef struct {
	uint32_t *cell;
	int32_t *energy;
	uint8_t *type;
	uint64_t n, cap;
	int over;        /* an agent's energy exceeded the limit */
} Part;

typedef struct {
	ppo_params p;
	uint32_t key[2];
	uint64_t cap, n, agent_steps;
	uint32_t step;
	int T;
	uint32_t *band_of_row;   /* band of each grid row */
	uint32_t *band_c0;       /* first cell of each band, T + 1 entries */
	/* state */
	uint32_t *cell;
	int32_t *energy;
	uint8_t *type;
	uint32_t *grass;
	/* scratch, per agent id */
	uint32_t *mcell;
	int32_t *menergy;
	uint8_t *mdice;
	uint8_t *killed;
	int32_t *child_e;
	uint32_t *part;          /* survivors grouped by band, id order */
	uint32_t *sorted;        /* survivors sorted by cell */
	uint32_t *ccount;        /* per-cell counts, C + T */
	uint64_t *tb;            /* [T][T] survivors of thread t for band b */
	uint64_t *bstart;        /* T + 1 */
	Part *out;               /* per band */
	uint64_t *ostart;        /* T + 1 */
	ppo_stats *partial;      /* per thread */
} Sim;

static Sim *sim_create(const ppo_params *p, uint64_t seed, uint64_t cap, int T) {
	const uint64_t n0 = (uint64_t)p->init_sheep + p->init_wolves;
	if (!p->grid_x || !p->grid_y || !p->grass_restart || cap < n0) return NULL;
	if (T < 1) T = 1;
	if ((uint32_t)T > p->grid_y) T = (int)p->grid_y;
	Sim *s = calloc(1, sizeof(Sim));
	s->p = *p;
	s->p.grid_xy = p->grid_x * p->grid_y;
	s->key[0] = (uint32_t)seed;
	s->key[1] = (uint32_t)(seed >> 32);
	s->cap = cap;
	s->T = T;
	const uint64_t C = s->p.grid_xy;
	s->band_of_row = xalloc(p->grid_y * sizeof(uint32_t));
	s->band_c0 = xalloc((T + 1) * sizeof(uint32_t));
	for (int b = 0; b <= T; b++) {
		const uint32_t y0 = (uint32_t)((uint64_t)p->grid_y * b / T);
		s->band_c0[b] = y0 * p->grid_x;
		if (b < T) {
			const uint32_t y1 = (uint32_t)((uint64_t)p->grid_y * (b + 1) / T);
			for (uint32_t y = y0; y < y1; y++) s->band_of_row[y] = (uint32_t)b;
		}
	}
	s->cell = xalloc(cap * 4);
	s->energy = xalloc(cap * 4);
	s->type = xalloc(cap);
	s->grass = xalloc(C * 4);
	s->mcell = xalloc(cap * 4);
	s->menergy = xalloc(cap * 4);
	s->mdice = xalloc(cap);
	s->killed = xalloc(cap);
	s->child_e = xalloc(cap * 4);
	s->part = xalloc(cap * 4);
	s->sorted = xalloc(cap * 4);
	s->ccount = xalloc((C + T) * 4);
	s->tb = xalloc((size_t)T * T * 8);
	s->bstart = xalloc((T + 1) * 8);
	s->out = calloc((size_t)T, sizeof(Part));
	s->ostart = xalloc((T + 1) * 8);
	s->partial = xalloc((size_t)T * sizeof(ppo_stats));

	/* init of cells and agents, as the oracle (pp_gpu.cl:368-439) */
#pragma omp parallel for schedule(static) num_threads(T)
	for (uint64_t c = 0; c < C; c++) {
		uint32_t r[4];
		philox(s->key, (uint32_t)c, 0, S_INIT_CELL, 0, r);
		s->grass[c] = bounded(r[0], 2) == 1 ? 0 : 1 + bounded(r[1], p->grass_restart);
	}
#pragma omp parallel for schedule(static) num_threads(T)
	for (uint64_t i = 0; i < n0; i++) {
		uint32_t r[4];
		const int wolf = i >= p->init_sheep;
		const uint32_t gain = wolf ? p->wolves_gain_from_food : p->sheep_gain_from_food;
		philox(s->key, (uint32_t)i, 0, S_INIT_AGENT, 0, r);
		s->cell[i] = bounded(r[1], p->grid_y) * p->grid_x + bounded(r[0], p->grid_x);
		s->energy[i] = (int32_t)(1 + bounded(r[2], 2 * gain));
		s->type[i] = (uint8_t)wolf;
	}
	s->n = n0;
	return s;
}

static void sim_destroy(Sim *s) {
	if (!s) return;
	for (int b = 0; b < s->T; b++) {
		free(s->out[b].cell);
		free(s->out[b].energy);
		free(s->out[b].type);
	}
	free(s->out); free(s->ostart); free(s->partial);
	free(s->band_of_row); free(s->band_c0);
	free(s->cell); free(s->energy); free(s->type); free(s->grass);
	free(s->mcell); free(s->menergy); free(s->mdice); free(s->killed);
	free(s->child_e); free(s->part); free(s->sorted); free(s->ccount);
	free(s->tb); free(s->bstart);
	free(s);
}

static void part_reserve(Part *o, uint64_t need) {
	if (need <= o->cap) return;
	const uint64_t cap = need + need / 4 + 64;
	free(o->cell); free(o->energy); free(o->type);
	o->cell = xalloc(cap * 4);
	o->energy = xalloc(cap * 4);
	o->type = xalloc(cap);
	o->cap = cap;
}

static inline uint32_t moved(const ppo_params *p, uint32_t cell, uint32_t dir) {
	/* the reference GPU table (pp_gpu.cl:780-803), torus wrap */
	uint32_t x = cell % p->grid_x, y = cell / p->grid_x;
	if (dir == 1) x = x + 1 == p->grid_x ? 0 : x + 1;
	else if (dir == 2) x = x ? x - 1 : p->grid_x - 1;
	else if (dir == 3) y = y + 1 == p->grid_y ? 0 : y + 1;
	else if (dir == 4) y = y ? y - 1 : p->grid_y - 1;
	return y * p->grid_x + x;
}

/* Statistics of the current state (pp_common.cl:55-78 fields, exact sums). */
static void sim_stats(Sim *s, ppo_stats *row) {
	const int T = s->T;
	const uint64_t n = s->n, C = s->p.grid_xy;
#pragma omp parallel num_threads(T)
	{
		const int t = omp_get_thread_num();
		ppo_stats a;
		memset(&a, 0, sizeof(a));
		for (uint64_t i = n * t / T; i < n * (t + 1) / T; i++) {
			if (s->type[i]) { a.wolves++; a.wolves_en += (uint64_t)s->energy[i]; }
			else { a.sheep++; a.sheep_en += (uint64_t)s->energy[i]; }
		}
		for (uint64_t c = C * t / T; c < C * (t + 1) / T; c++) {
			a.grass += s->grass[c] == 0;
			a.grass_en += s->grass[c];
		}
		s->partial[t] = a;
	}
	memset(row, 0, sizeof(*row));
	for (int t = 0; t < T; t++) {
		row->sheep += s->partial[t].sheep;
		row->wolves += s->partial[t].wolves;
		row->grass += s->partial[t].grass;
		row->sheep_en += s->partial[t].sheep_en;
		row->wolves_en += s->partial[t].wolves_en;
		row->grass_en += s->partial[t].grass_en;
	}
}

/* One iteration; 0, or -8 when the agents outgrow the capacity. */
static int sim_step(Sim *s) {
	const ppo_params *p = &s->p;
	const int T = s->T;
	const uint32_t st = s->step;
	const uint64_t n = s->n;
	int rc = 0;
#pragma omp parallel num_threads(T)
	{
		const int t = omp_get_thread_num();
		uint64_t *cnt = s->tb + (size_t)t * T;
		/* 1. move a run of ids, count survivors per destination band */
		for (int b = 0; b < T; b++) cnt[b] = 0;
		const uint64_t i0 = n * t / T, i1 = n * (t + 1) / T;
		for (uint64_t i = i0; i < i1; i++) {
			uint32_t r[4];
			philox(s->key, (uint32_t)i, st, S_MOVE, 0, r);
			s->mdice[i] = (uint8_t)bounded(r[1], 100);
			s->menergy[i] = s->energy[i] - 1;
			s->mcell[i] = moved(p, s->cell[i], bounded(r[0], 5));
			if (s->menergy[i] >= 1) cnt[s->band_of_row[s->mcell[i] / p->grid_x]]++;
		}
#pragma omp barrier
#pragma omp single
		{
			/* band starts, and each thread's offset inside each band */
			uint64_t run = 0;
			for (int b = 0; b < T; b++) {
				s->bstart[b] = run;
				for (int u = 0; u < T; u++) {
					const uint64_t c = s->tb[(size_t)u * T + b];
					s->tb[(size_t)u * T + b] = run;
					run += c;
				}
			}
			s->bstart[T] = run;
		}
		/* 2. scatter survivors to their bands, id order kept */
		for (uint64_t i = i0; i < i1; i++)
			if (s->menergy[i] >= 1)
				s->part[cnt[s->band_of_row[s->mcell[i] / p->grid_x]]++] = (uint32_t)i;
#pragma omp barrier
		/* 3. band t: grass, counting sort by cell, actions */
		const uint32_t c0 = s->band_c0[t], c1 = s->band_c0[t + 1];
		for (uint32_t c = c0; c < c1; c++)
			if (s->grass[c]) s->grass[c]--;
		uint32_t *cc = s->ccount + c0 + t;   /* c1 - c0 + 1 counters */
		memset(cc, 0, (size_t)(c1 - c0 + 1) * 4);
		const uint64_t b0 = s->bstart[t], b1 = s->bstart[t + 1];
		for (uint64_t q = b0; q < b1; q++) cc[s->mcell[s->part[q]] - c0 + 1]++;
		for (uint32_t k = 0; k < c1 - c0; k++) cc[k + 1] += cc[k];
		uint32_t *srt = s->sorted + b0;
		for (uint64_t q = b0; q < b1; q++) {
			const uint32_t id = s->part[q];
			srt[cc[s->mcell[id] - c0]++] = id;
		}
		Part *o = &s->out[t];
		part_reserve(o, 2 * (b1 - b0));
		o->over = 0;
		uint64_t no = 0;
		for (uint64_t q0 = 0; q0 < b1 - b0;) {
			const uint32_t c = s->mcell[srt[q0]];
			uint64_t q1 = q0 + 1;
			while (q1 < b1 - b0 && s->mcell[srt[q1]] == c) q1++;
			uint32_t *A = srt + q0;
			const uint32_t k = (uint32_t)(q1 - q0);
			/* Fisher-Yates from the last position; word m % 4 of block
			 * m / 4 of the (cell, step) shuffle stream, m = k - 1 - i */
			uint32_t r[4];
			for (uint32_t i = k - 1, m = 0; i > 0; i--, m++) {
				if ((m & 3u) == 0) philox(s->key, c, st, S_SHUFFLE, m >> 2, r);
				const uint32_t j = bounded(r[m & 3u], i + 1);
				const uint32_t tmp = A[i];
				A[i] = A[j];
				A[j] = tmp;
			}
			for (uint32_t a = 0; a < k; a++) {
				s->killed[A[a]] = 0;
				s->child_e[A[a]] = -1;
			}
			for (uint32_t a = 0; a < k; a++) {
				const uint32_t id = A[a];
				if (s->killed[id]) continue;
				int32_t e = s->menergy[id];
				uint32_t thr, prob;
				if (s->type[id] == 0) {          /* sheep: live grass */
					if (s->grass[c] == 0) {
						s->grass[c] = p->grass_restart;
						e += (int32_t)p->sheep_gain_from_food;
					}
					thr = p->sheep_reproduce_threshold;
					prob = p->sheep_reproduce_prob;
				} else {                         /* wolf: first live sheep */
					for (uint32_t b = 0; b < k; b++) {
						const uint32_t v = A[b];
						if (s->type[v] == 0 && !s->killed[v]) {
							s->killed[v] = 1;
							e += (int32_t)p->wolves_gain_from_food;
							break;
						}
					}
					thr = p->wolves_reproduce_threshold;
					prob = p->wolves_reproduce_prob;
				}
				if (e > (int32_t)thr && s->mdice[id] < prob) {
					const int32_t ce = e / 2;
					e -= ce;
					s->child_e[id] = ce;
				}
				/* the oracle's energy limit (2^24 - 1 after the action) */
				o->over |= e > 0x00FFFFFF;
				s->menergy[id] = e;
			}
			for (uint32_t a = 0; a < k; a++) {
				const uint32_t id = A[a];
				if (!s->killed[id]) {
					o->cell[no] = c;
					o->energy[no] = s->menergy[id];
					o->type[no] = s->type[id];
					no++;
				}
				if (s->child_e[id] >= 0) {
					o->cell[no] = c;
					o->energy[no] = s->child_e[id];
					o->type[no] = s->type[id];
					no++;
				}
			}
			q0 = q1;
		}
		o->n = no;
#pragma omp barrier
#pragma omp single
		{
			uint64_t run = 0;
			for (int b = 0; b < T; b++) {
				s->ostart[b] = run;
				run += s->out[b].n;
			}
			s->ostart[T] = run;
			int over = 0;
			for (int b = 0; b < T; b++) over |= s->out[b].over;
			if (over) rc = -9;
			else if (run > s->cap) rc = -8;
		}
		/* 4. concatenate the bands' parts in band (= cell) order */
		if (!rc) {
			const uint64_t at = s->ostart[t];
			memcpy(s->cell + at, o->cell, o->n * 4);
			memcpy(s->energy + at, o->energy, o->n * 4);
			memcpy(s->type + at, o->type, o->n);
		}
	}
	if (rc) return rc;
	s->agent_steps += n;
	s->n = s->ostart[T];
	s->step++;
	return 0;
}

static double now_s(void) {
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
	const char *params = "config.txt", *stats = "stats.txt", *raw = NULL;
	uint64_t seed = 0, cap = 0;
	long iters_override = -1;
	int quiet = 0, threads = 0, opt;
	while ((opt = getopt(argc, argv, "p:s:r:m:n:t:qb:")) != -1) {
		switch (opt) {
		case 'p': params = optarg; break;
		case 's': stats = optarg; break;
		case 'r': seed = strtoull(optarg, NULL, 0); break;
		case 'm': cap = strtoull(optarg, NULL, 0); break;
		case 'n': iters_override = strtol(optarg, NULL, 0); break;
		case 't': threads = atoi(optarg); break;
		case 'q': quiet = 1; break;
		case 'b': raw = optarg; break;
		default: return -1;
		}
	}
	ppo_params p;
	int rc = ppo_load_params(params, &p);
	if (rc) { fprintf(stderr, "Error: parameters file (%d)\n", rc); return rc; }
	if (iters_override >= 0) p.iters = (uint32_t)iters_override;
	const uint64_t n0 = (uint64_t)p.init_sheep + p.init_wolves;
	if (!cap) cap = n0 * 4 > 16777216 ? n0 * 4 : 16777216;
	if (threads <= 0) threads = omp_get_max_threads();
	Sim *s = sim_create(&p, seed, cap, threads);
	if (!s) { fprintf(stderr, "Error: invalid arguments\n"); return -2; }
	ppo_stats *rows = calloc((size_t)p.iters + 1, sizeof(ppo_stats));
	sim_stats(s, &rows[0]);
	const double t0 = now_s();
	uint32_t t;
	for (t = 0; t < p.iters; t++) {
		rc = sim_step(s);
		if (rc) break;
		sim_stats(s, &rows[t + 1]);
	}
	const double t1 = now_s();
	if (rc) fprintf(stderr, "Error: %s at step %u\n",
		rc == -9 ? "energy above its limit" : "out of resources", t);
	printf("agent_steps=%llu seconds=%.6f agent_steps_per_s=%.6e threads=%d\n",
		(unsigned long long)s->agent_steps, t1 - t0,
		(double)s->agent_steps / (t1 - t0), s->T);
	if (raw) {
		FILE *fp = fopen(raw, "wb");
		if (!fp || fwrite(rows, sizeof(ppo_stats), (size_t)t + 1, fp) != (size_t)t + 1)
			rc = rc ? rc : -6;
		if (fp) fclose(fp);
	}
	if (!quiet && !rc) rc = ppo_stats_save(stats, rows, p.iters + 1, p.grid_xy);
	free(rows);
	sim_destroy(s);
	return rc;
}
