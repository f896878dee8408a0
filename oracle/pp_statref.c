/*
 * pp_statref.c -- independent CPU simulation of the PPHPC model used as the
 * "reference side" of the statistical-equivalence tests.
 *
 * TEST INFRASTRUCTURE ONLY. It shares nothing with the oracle or the GPU
 * implementation except the model rules: it follows the canonical NetLogo
 * specification (netlogo/pphpc.nlogo:18-159) step by step, with its own
 * random number generator and its own orderings:
 *   - RNG: MT19937 (Matsumoto & Nishimura 1998, init_genrand seeding), bounded
 *     draws by rejection (uniform, unlike the multiply-high map of the step
 *     under test);
 *   - go (pphpc.nlogo:63-98): all turtles move (random 5: directions 0..3 are
 *     headings 0/90/180/270 and fd 1, 4 stays; pphpc.nlogo:116-122), lose one
 *     energy and die below 1; all patches grow grass (pphpc.nlogo:151-159);
 *     then ALL turtles act in one global random order (ask turtles): sheep eat
 *     green grass (pphpc.nlogo:124-131), wolves catch ONE RANDOM sheep of the
 *     patch that is not newborn (who <= maxwho, pphpc.nlogo:144-149); both try
 *     to reproduce (pphpc.nlogo:133-142); newborns are not asked this tick;
 *   - setup (pphpc.nlogo:18-60): patches alive with probability 1/2, else
 *     countdown 1 + random R; sheep then wolves at random patches with energy
 *     1 + random (2 * gain).
 * The step under test orders actions per cell (Java Cell.agentActions,
 * Cell.java:236-259) and lets a wolf take the first live sheep of the
 * shuffled cell list (Wolf.java:66-92); both are equal in distribution to the
 * NetLogo choices restated here, which is what the statistical tests check.
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

/* MT19937 output of init_genrand(seed), for the known-answer test. */
void ppr_mt_outputs(uint32_t seed, uint32_t n, uint32_t *out) {
	mt19937 g;
	mt_seed(&g, seed);
	for (uint32_t i = 0; i < n; i++) out[i] = mt_next(&g);
}
