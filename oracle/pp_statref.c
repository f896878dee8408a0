/*
 * pp_statref.c -- independent CPU simulations of the PPHPC model used as the
 * "reference side" of the statistical-equivalence tests.
 *
 * TEST INFRASTRUCTURE ONLY. They share nothing with the oracle or the GPU
 * implementation except the model rules, and use their own random number
 * generator and their own orderings: MT19937 (Matsumoto & Nishimura 1998,
 * init_genrand seeding), bounded draws by rejection (uniform, unlike the
 * multiply-high map of the step under test). Two rule sets:
 *
 * ppr_run_cells -- the Java single-thread semantics (SimWorker.java:183-232,
 *   Cell.java:175-308, AbstractAgent.java:91-167, Sheep.java:66-83,
 *   Wolf.java:66-92), the semantics the step under test follows:
 *   - move: cells in index order, each cell's agents in list order lose one
 *     energy, those still above 0 draw one of 5 moves (stay, up, right,
 *     down, left) and are appended to the destination cell's list;
 *     then every positive grass countdown drops by 1;
 *   - act: cells in index order; the cell's list is shuffled (Fisher-Yates
 *     from the last position, as Collections.shuffle); each live agent in
 *     that order eats (a sheep the live grass; a wolf the FIRST live sheep
 *     of the shuffled list, possibly one that has already acted) and, if its
 *     energy exceeds the threshold, draws U[0,100) against the reproduction
 *     probability; a child gets floor(e/2) and joins the cell's newborns;
 *   - next state per cell: the live agents in shuffled order, then the
 *     newborns. The grid is a torus in both directions (the Java space wraps
 *     its right/left neighbours across rows, VonNeumann2DTorusSpace.java:
 *     73-76; NetLogo and the OpenCL kernels use a torus, as here).
 *
 * ppr_run -- the NetLogo rules (netlogo/pphpc.nlogo:18-159): all turtles
 *   move (random 5: headings 0/90/180/270 and fd 1, 4 stays;
 *   pphpc.nlogo:116-122), lose one energy and die below 1; all patches grow
 *   grass (pphpc.nlogo:151-159); then ALL turtles act in one global random
 *   order (ask turtles): sheep eat green grass (pphpc.nlogo:124-131), wolves
 *   catch ONE RANDOM sheep of the patch that is not newborn (who <= maxwho,
 *   pphpc.nlogo:144-149); both try to reproduce (pphpc.nlogo:133-142);
 *   newborns are not asked this tick.
 *
 * The two differ in the wolf's choice of prey: the Java wolf takes the first
 * live sheep of the shuffled list, so a sheep that acted before the wolf is
 * a likelier victim than one still to act; the NetLogo wolf draws uniformly.
 * Both share setup (pphpc.nlogo:18-60, CellGrassInitCoinRandCounter.java:
 * 49-67): patches alive with probability 1/2, else countdown 1 + random R;
 * sheep then wolves at random patches with energy 1 + random (2 * gain).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- MT19937 */

typedef struct {
	uint32_t mt[624];
	int mti;
} mt19937;

static void mt_seed(mt19937 *g, uint32_t s) {
	g->mt[0] = s;
	for (int i = 1; i < 624; i++)
		g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
	g->mti = 624;
}

static uint32_t mt_next(mt19937 *g) {
	if (g->mti >= 624) {
		for (int i = 0; i < 624; i++) {
			uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
			g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
		}
		g->mti = 0;
	}
	uint32_t y = g->mt[g->mti++];
	y ^= y >> 11;
	y ^= (y << 7) & 0x9d2c5680u;
	y ^= (y << 15) & 0xefc60000u;
	y ^= y >> 18;
	return y;
}

/* Uniform in [0, n) by rejection of the biased tail. */
static uint32_t mt_below(mt19937 *g, uint32_t n) {
	if (n <= 1) return 0;
	const uint32_t limit = (uint32_t)(0x100000000ull - (0x100000000ull % n));
	uint32_t x;
	do {
		x = mt_next(g);
	} while (limit != 0 && x >= limit);
	return x % n;
}

/* ------------------------------------------------------------------ model */

typedef struct {
	uint32_t init_sheep, sheep_gain, sheep_thr, sheep_prob;
	uint32_t init_wolves, wolves_gain, wolves_thr, wolves_prob;
	uint32_t grass_restart, gx, gy;
} ppr_params;

typedef struct {
	int32_t x, y, energy;
	uint8_t wolf, alive;
	uint64_t who;
} turtle;

typedef struct {
	ppr_params p;
	mt19937 rng;
	turtle *t;
	size_t n, cap;
	uint64_t next_who, maxwho;
	int32_t *countdown;
	/* per-patch lists of sheep indices, rebuilt after the move */
	uint32_t *head, *next;
	uint32_t *order;
	size_t order_cap;
	uint32_t *cand;
	size_t cand_cap;
} ppr_world;

static void *xrealloc(void *ptr, size_t sz) {
	void *q = realloc(ptr, sz ? sz : 1);
	if (!q) { fprintf(stderr, "pp_statref: out of memory\n"); abort(); }
	return q;
}

static void push_turtle(ppr_world *w, turtle tu) {
	if (w->n == w->cap) {
		w->cap = w->cap ? 2 * w->cap : 1024;
		w->t = xrealloc(w->t, w->cap * sizeof(turtle));
	}
	w->t[w->n++] = tu;
}

static void gather(const ppr_world *w, uint64_t out[7]) {
	/* gather-stats, pphpc.nlogo:100-114 (sums instead of means) */
	memset(out, 0, 7 * sizeof(uint64_t));
	for (size_t i = 0; i < w->n; i++) {
		if (!w->t[i].alive) continue;
		if (w->t[i].wolf) { out[1]++; out[4] += (uint64_t)w->t[i].energy; }
		else { out[0]++; out[3] += (uint64_t)w->t[i].energy; }
	}
	const size_t C = (size_t)w->p.gx * w->p.gy;
	for (size_t c = 0; c < C; c++) {
		if (w->countdown[c] <= 0) out[2]++;
		else out[5] += (uint64_t)w->countdown[c];
	}
}

static void update_maxwho(ppr_world *w) {
	uint64_t m = 0;
	for (size_t i = 0; i < w->n; i++)
		if (w->t[i].alive && !w->t[i].wolf && w->t[i].who > m) m = w->t[i].who;
	w->maxwho = m;
}

static void setup(ppr_world *w, const ppr_params *p, uint32_t seed) {
	memset(w, 0, sizeof(*w));
	w->p = *p;
	mt_seed(&w->rng, seed);
	const size_t C = (size_t)p->gx * p->gy;
	w->countdown = xrealloc(NULL, C * sizeof(int32_t));
	w->head = xrealloc(NULL, C * sizeof(uint32_t));
	for (size_t c = 0; c < C; c++)
		w->countdown[c] = mt_below(&w->rng, 2) == 1
			? 1 + (int32_t)mt_below(&w->rng, p->grass_restart) : 0;
	for (uint32_t i = 0; i < p->init_sheep + p->init_wolves; i++) {
		turtle tu;
		tu.wolf = i >= p->init_sheep;
		uint32_t gain = tu.wolf ? p->wolves_gain : p->sheep_gain;
		tu.energy = 1 + (int32_t)mt_below(&w->rng, 2 * gain);
		tu.x = (int32_t)mt_below(&w->rng, p->gx);
		tu.y = (int32_t)mt_below(&w->rng, p->gy);
		tu.alive = 1;
		tu.who = w->next_who++;
		push_turtle(w, tu);
	}
	update_maxwho(w);
}

static void go(ppr_world *w) {
	const ppr_params *p = &w->p;
	const size_t C = (size_t)p->gx * p->gy;
	/* 1. move, lose energy, die (pphpc.nlogo:66-70, 116-122) */
	for (size_t i = 0; i < w->n; i++) {
		turtle *tu = &w->t[i];
		uint32_t d = mt_below(&w->rng, 5);
		if (d == 0) tu->y = (tu->y + 1) % (int32_t)p->gy;
		else if (d == 1) tu->x = (tu->x + 1) % (int32_t)p->gx;
		else if (d == 2) tu->y = (tu->y + (int32_t)p->gy - 1) % (int32_t)p->gy;
		else if (d == 3) tu->x = (tu->x + (int32_t)p->gx - 1) % (int32_t)p->gx;
		tu->energy -= 1;
		if (tu->energy < 1) tu->alive = 0;
	}
	/* compact the dead */
	size_t k = 0;
	for (size_t i = 0; i < w->n; i++)
		if (w->t[i].alive) w->t[k++] = w->t[i];
	w->n = k;
	/* 2. grow grass (pphpc.nlogo:151-159) */
	for (size_t c = 0; c < C; c++)
		if (w->countdown[c] > 0) w->countdown[c]--;
	/* sheep lists per patch */
	w->next = xrealloc(w->next, (w->n ? w->n : 1) * sizeof(uint32_t));
	for (size_t c = 0; c < C; c++) w->head[c] = UINT32_MAX;
	for (size_t i = 0; i < w->n; i++) {
		if (w->t[i].wolf) continue;
		size_t c = (size_t)w->t[i].y * p->gx + (size_t)w->t[i].x;
		w->next[i] = w->head[c];
		w->head[c] = (uint32_t)i;
	}
	/* 3. ask turtles in random order (the ones existing now) */
	const size_t n0 = w->n;
	if (w->order_cap < n0) {
		w->order_cap = n0;
		w->order = xrealloc(w->order, n0 * sizeof(uint32_t));
	}
	for (size_t i = 0; i < n0; i++) w->order[i] = (uint32_t)i;
	for (size_t i = n0; i > 1; i--) {
		size_t j = mt_below(&w->rng, (uint32_t)i);
		uint32_t tmp = w->order[i - 1];
		w->order[i - 1] = w->order[j];
		w->order[j] = tmp;
	}
	for (size_t o = 0; o < n0; o++) {
		const uint32_t id = w->order[o];
		turtle *tu = &w->t[id];
		if (!tu->alive) continue;
		size_t c = (size_t)tu->y * p->gx + (size_t)tu->x;
		uint32_t thr, prob;
		if (!tu->wolf) {
			if (w->countdown[c] <= 0) {
				w->countdown[c] = (int32_t)p->grass_restart;
				tu->energy += (int32_t)p->sheep_gain;
			}
			thr = p->sheep_thr;
			prob = p->sheep_prob;
		} else {
			/* one-of sheep-here with [who <= maxwho] */
			size_t nc = 0;
			for (uint32_t s = w->head[c]; s != UINT32_MAX; s = w->next[s]) {
				if (w->t[s].alive && w->t[s].who <= w->maxwho) {
					if (nc == w->cand_cap) {
						w->cand_cap = w->cand_cap ? 2 * w->cand_cap : 64;
						w->cand = xrealloc(w->cand, w->cand_cap * sizeof(uint32_t));
					}
					w->cand[nc++] = s;
				}
			}
			if (nc) {
				uint32_t prey = w->cand[mt_below(&w->rng, (uint32_t)nc)];
				w->t[prey].alive = 0;
				tu = &w->t[id];
				tu->energy += (int32_t)p->wolves_gain;
			}
			thr = p->wolves_thr;
			prob = p->wolves_prob;
		}
		if (tu->energy > (int32_t)thr && mt_below(&w->rng, 100) < prob) {
			int32_t child = tu->energy / 2;
			tu->energy -= child;
			turtle nb = *tu;
			nb.energy = child;
			nb.who = w->next_who++;
			push_turtle(w, nb);     /* may move w->t */
		}
	}
	k = 0;
	for (size_t i = 0; i < w->n; i++)
		if (w->t[i].alive) w->t[k++] = w->t[i];
	w->n = k;
	update_maxwho(w);
}

static void destroy(ppr_world *w) {
	free(w->t); free(w->countdown); free(w->head); free(w->next);
	free(w->order); free(w->cand);
}

/* Runs setup + iters ticks; writes (iters + 1) x 7 rows of sheep, wolves,
 * grass alive, sheep energy, wolf energy, countdown sum, 0. Returns 0. */
int ppr_run(const uint32_t params[11], uint32_t seed, uint32_t iters,
	uint64_t *rows) {
	ppr_params p;
	memcpy(&p, params, sizeof(p));
	if (p.gx == 0 || p.gy == 0 || p.grass_restart == 0) return -2;
	ppr_world w;
	setup(&w, &p, seed);
	gather(&w, rows);
	for (uint32_t t = 1; t <= iters; t++) {
		go(&w);
		gather(&w, rows + 7 * (size_t)t);
	}
	destroy(&w);
	return 0;
}

/* ------------------------------------------------- Java single-thread rules */

typedef struct {
	int32_t energy;
	uint8_t wolf;
} jagent;

static void *xalloc(size_t n, size_t sz) { return xrealloc(NULL, n * sz); }

/* Runs setup + iters steps of the Java single-thread semantics; rows as
 * ppr_run. The state is kept in cell order: agents of cell 0, then cell 1,
 * ..., each cell's agents in its list order (cnt[c] of them). */
int ppr_run_cells(const uint32_t params[11], uint32_t seed, uint32_t iters,
	uint64_t *rows) {
	ppr_params p;
	memcpy(&p, params, sizeof(p));
	if (p.gx == 0 || p.gy == 0 || p.grass_restart == 0) return -2;
	const size_t C = (size_t)p.gx * p.gy;
	mt19937 rng;
	mt_seed(&rng, seed);
	int32_t *grass = xalloc(C, sizeof(int32_t));
	for (size_t c = 0; c < C; c++)
		grass[c] = mt_below(&rng, 2) == 1
			? 1 + (int32_t)mt_below(&rng, p.grass_restart) : 0;
	/* initial agents in creation order per cell */
	size_t n = (size_t)p.init_sheep + p.init_wolves, cap = n ? 2 * n : 16;
	jagent *ag = xalloc(cap, sizeof(jagent));
	uint32_t *acell = xalloc(cap, sizeof(uint32_t));
	for (size_t i = 0; i < n; i++) {
		const int wolf = i >= p.init_sheep;
		const uint32_t gain = wolf ? p.wolves_gain : p.sheep_gain;
		ag[i].energy = 1 + (int32_t)mt_below(&rng, 2 * gain);
		ag[i].wolf = (uint8_t)wolf;
		const uint32_t x = mt_below(&rng, p.gx), y = mt_below(&rng, p.gy);
		acell[i] = y * p.gx + x;
	}
	uint32_t *cnt = xalloc(C + 1, sizeof(uint32_t));
	uint32_t *start = xalloc(C + 1, sizeof(uint32_t));
	/* order the initial agents by cell, keeping creation order (the cell
	 * lists of putInitAgent) */
	jagent *cur = xalloc(cap, sizeof(jagent));
	memset(cnt, 0, (C + 1) * sizeof(uint32_t));
	for (size_t i = 0; i < n; i++) cnt[acell[i]]++;
	start[0] = 0;
	for (size_t c = 0; c < C; c++) start[c + 1] = start[c] + cnt[c];
	for (size_t i = 0; i < n; i++) cur[start[acell[i]]++] = ag[i];
	start[0] = 0;
	for (size_t c = 0; c < C; c++) start[c + 1] = start[c] + cnt[c];
	jagent *mov = NULL, *kids = NULL, *nxt = NULL;
	uint32_t *mcell = NULL, *mcnt = xalloc(C + 1, sizeof(uint32_t));
	size_t mcap = 0, kcap = 0, ncap = 0;

	for (uint32_t t = 0; t <= iters; t++) {
		/* statistics of the state (cell order) */
		uint64_t *row = rows + 7 * (size_t)t;
		memset(row, 0, 7 * sizeof(uint64_t));
		for (size_t i = 0; i < n; i++) {
			if (cur[i].wolf) { row[1]++; row[4] += (uint64_t)cur[i].energy; }
			else { row[0]++; row[3] += (uint64_t)cur[i].energy; }
		}
		for (size_t c = 0; c < C; c++) {
			if (grass[c] <= 0) row[2]++;
			else row[5] += (uint64_t)grass[c];
		}
		if (t == iters) break;
		/* 1. move, in cell order and list order */
		if (mcap < n) {
			mcap = n;
			mov = xrealloc(mov, mcap * sizeof(jagent));
			mcell = xrealloc(mcell, mcap * sizeof(uint32_t));
		}
		size_t m = 0;
		for (size_t c = 0; c < C; c++) {
			const uint32_t x = (uint32_t)(c % p.gx), y = (uint32_t)(c / p.gx);
			for (uint32_t i = start[c]; i < start[c + 1]; i++) {
				jagent a = cur[i];
				a.energy -= 1;
				if (a.energy <= 0) continue;
				uint32_t nx = x, ny = y;
				switch (mt_below(&rng, 5)) {
				case 1: ny = (y + p.gy - 1) % p.gy; break;   /* up */
				case 2: nx = (x + 1) % p.gx; break;          /* right */
				case 3: ny = (y + 1) % p.gy; break;          /* down */
				case 4: nx = (x + p.gx - 1) % p.gx; break;   /* left */
				default: break;                               /* stay */
				}
				mov[m] = a;
				mcell[m] = ny * p.gx + nx;
				m++;
			}
		}
		/* 2. grass */
		for (size_t c = 0; c < C; c++)
			if (grass[c] > 0) grass[c]--;
		/* destination lists in append order (stable by cell) */
		memset(mcnt, 0, (C + 1) * sizeof(uint32_t));
		for (size_t i = 0; i < m; i++) mcnt[mcell[i]]++;
		uint32_t run = 0;
		for (size_t c = 0; c < C; c++) {
			const uint32_t k = mcnt[c];
			mcnt[c] = run;
			run += k;
		}
		mcnt[C] = run;
		if (ncap < m) {
			ncap = m;
			nxt = xrealloc(nxt, ncap * sizeof(jagent));
		}
		for (size_t i = 0; i < m; i++) nxt[mcnt[mcell[i]]++] = mov[i];
		for (size_t c = C; c > 0; c--) mcnt[c] = mcnt[c - 1];
		mcnt[0] = 0;
		/* 3. actions per cell; the next state is written in cell order */
		size_t out = 0;
		if (cap < 2 * m + 1) {
			cap = 2 * m + 1;
			cur = xrealloc(cur, cap * sizeof(jagent));
		}
		for (size_t c = 0; c < C; c++) {
			jagent *L = nxt + mcnt[c];
			const uint32_t k = mcnt[c + 1] - mcnt[c];
			for (uint32_t i = k; i > 1; i--) {
				const uint32_t j = mt_below(&rng, i);
				const jagent tmp = L[i - 1];
				L[i - 1] = L[j];
				L[j] = tmp;
			}
			if (kcap < k) {
				kcap = k;
				kids = xrealloc(kids, kcap * sizeof(jagent));
			}
			uint32_t nk = 0;
			for (uint32_t i = 0; i < k; i++) {
				jagent *a = &L[i];
				if (a->energy <= 0) continue;
				uint32_t thr, prob;
				if (!a->wolf) {
					if (grass[c] <= 0) {
						grass[c] = (int32_t)p.grass_restart;
						a->energy += (int32_t)p.sheep_gain;
					}
					thr = p.sheep_thr;
					prob = p.sheep_prob;
				} else {
					for (uint32_t o = 0; o < k; o++) {
						if (!L[o].wolf && L[o].energy > 0) {
							L[o].energy = 0;
							a->energy += (int32_t)p.wolves_gain;
							break;
						}
					}
					thr = p.wolves_thr;
					prob = p.wolves_prob;
				}
				if (a->energy > (int32_t)thr && mt_below(&rng, 100) < prob) {
					jagent child = *a;
					child.energy = a->energy / 2;
					a->energy -= child.energy;
					kids[nk++] = child;
				}
			}
			start[c] = (uint32_t)out;
			for (uint32_t i = 0; i < k; i++)
				if (L[i].energy > 0) cur[out++] = L[i];
			for (uint32_t i = 0; i < nk; i++) cur[out++] = kids[i];
		}
		start[C] = (uint32_t)out;
		n = out;
	}
	free(grass); free(ag); free(acell); free(cnt); free(start); free(cur);
	free(mov); free(mcell); free(mcnt); free(kids); free(nxt);
	return 0;
}

/* MT19937 output of init_genrand(seed), for the known-answer test. */
void ppr_mt_outputs(uint32_t seed, uint32_t n, uint32_t *out) {
	mt19937 g;
	mt_seed(&g, seed);
	for (uint32_t i = 0; i < n; i++) out[i] = mt_next(&g);
}
