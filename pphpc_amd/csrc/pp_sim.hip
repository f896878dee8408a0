/*
 * pp_sim.hip -- MI355X (gfx950) implementation of the PPHPC per-iteration step
 * behind the C ABI of include/pphpc.h.
 *
 * Replaces the hot loop of the reference GPU simulator, ppg_simulate
 * (opencl/pp/pp_gpu.c:671-935) and its kernels (opencl/pp/pp_gpu.cl:368-979):
 *
 *   reference                                this implementation
 *   ---------------------------------------  -----------------------------------
 *   init_cell, init_agent (pp_gpu.cl:368)    k_init_cells, k_init_agents
 *   grass (pp_gpu.cl:448)                    lazy regrow step per cell plus an
 *                                            O(1) ring update in k_finalize
 *   move_agent (pp_gpu.cl:744)               k_move (+ fused radix histograms)
 *   cl_ops sbitonic sort (pp_gpu.c:826)      k_sort_pass: LDS-staged onesweep
 *                                            LSD radix sort, 8-bit digits of the
 *                                            cell, stable (= key cell<<32|id)
 *   find_cell_idx (pp_gpu.cl:834)            segment heads found in k_action
 *   action_agent (pp_gpu.cl:883)             k_action: one lane per occupied
 *                                            cell, shuffled order, compaction
 *                                            and spawning by a block scan with
 *                                            decoupled look-back
 *   reduce_grass1/2, reduce_agent1/2         fused reductions (k_action,
 *   (pp_gpu.cl:485-736)                      k_finalize) with exact 64-bit sums
 *   stats readback every iteration           device ring of stats rows, one
 *   (pp_gpu.c:766, :840)                     copy per chunk of iterations
 *
 * All arithmetic is integer. The state, its ordering and every random draw are
 * defined so that a single-threaded CPU implementation reproduces the same
 * statistics bit for bit; see the header of each kernel.
 *
 * Device state of one simulation (capacity = A agents, C cells):
 *   agents (SoA)      cell u32[A], energy i32[A], type u8[A]       9 B/agent
 *   sort pairs        key u32[A], val u32[A], two buffers         16 B/agent
 *   grass             regrow step u32[C]                           4 B/cell
 *   grass ring        u64[R+1] cells per regrow step (mod R+1)
 *   look-back         u64 per (sort tile, digit) and per action tile
 *   control block     counters, live count, step, error, epoch
 *   stats ring        pp_stats64_t[4096]
 */
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "pphpc.h"
#include "pp_philox.h"

namespace {

constexpr int WG = 256;             /* threads per work-group (4 wave64) */
constexpr int WAVE = 64;
constexpr int NWAVE = WG / WAVE;
constexpr int RADIX = 256;          /* 8-bit digits */
constexpr int MAX_PASSES = 4;       /* 32-bit cell index */
constexpr int ACTION_SLOT = 7;      /* tile counter / epoch slot of k_action */
constexpr uint32_t DEAD_KEY = 0xFFFFFFFFu;
constexpr uint32_t STATS_RING = 4096;

/* Packed agent value carried through the sort and the action kernel. */
constexpr uint32_t V_ENERGY = 0x00FFFFFFu; /* energy, 24 bits */
constexpr uint32_t V_WOLF = 1u << 24;      /* type (0 sheep, 1 wolf) */
constexpr uint32_t V_REPRO = 1u << 25;     /* reproduction draw < probability */
constexpr uint32_t V_KILLED = 1u << 26;    /* eaten by a wolf this step */
constexpr uint32_t V_CHILD = 1u << 27;     /* reproduced this step */
constexpr uint32_t V_ODD = 1u << 28;       /* energy before the split was odd */

/* Look-back word: epoch (32) | flag (2) | value (30). */
constexpr uint64_t LB_AGG = 1;
constexpr uint64_t LB_INC = 2;

struct Ctl {
	uint32_t n;        /* agents of the current state */
	uint32_t step;     /* completed iterations */
	int32_t error;     /* 0 or a PP_* code; every kernel is a no-op once set */
	uint32_t err_step; /* iteration at which error was set */
	uint32_t live;     /* agents surviving the move of this iteration */
	uint32_t out_n;    /* agents after actions and births */
	uint32_t epoch;    /* look-back generation, +8 per iteration and init */
	uint32_t pad;
	uint32_t tile_ctr[8];
	unsigned long long dead_cells;  /* grass cells with countdown > 0 */
	unsigned long long cd_sum;      /* sum of grass countdowns */
	unsigned long long eaten;       /* grass eaten in this iteration */
	unsigned long long acc[4];      /* sheep, wolves, sheep energy, wolf energy */
	unsigned long long agent_steps; /* sum of n over completed iterations */
	unsigned long long occupied;    /* occupied cells in this iteration */
	unsigned long long sum_live;    /* sums over completed iterations of: */
	unsigned long long sum_out;     /*   live, out_n, occupied, eaten */
	unsigned long long sum_occ;
	unsigned long long sum_eaten;
};

struct Params {
	uint32_t key0, key1;
	uint32_t gx, gy, C;
	uint32_t R;
	uint32_t sheep_gain, wolves_gain;
	uint32_t sheep_thr, wolves_thr;
	uint32_t sheep_prob, wolves_prob;
	uint32_t init_sheep, n0;
	uint32_t cap;
	uint32_t passes;
};

/* ---------------------------------------------------------------- helpers */

__device__ __forceinline__ uint64_t lb_pack(uint32_t epoch, uint64_t flag,
	uint32_t v) {
	return ((uint64_t)epoch << 32) | (flag << 30) | (uint64_t)v;
}

__device__ __forceinline__ void lb_store(unsigned long long *p, uint64_t v) {
	__hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED,
		__HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lb_load(unsigned long long *p) {
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Exclusive prefix of the tiles before `tile` for one counter, by decoupled
 * look-back over words published by earlier tiles of the same epoch. Every
 * earlier tile was claimed (atomic counter) by a work-group that is already
 * running and publishes its aggregate before looking back itself, so the wait
 * always ends. */
__device__ __forceinline__ uint32_t lookback(unsigned long long *base,
	uint32_t stride, uint32_t tile, uint32_t epoch) {
	uint32_t excl = 0;
	int64_t t = (int64_t)tile - 1;
	while (t >= 0) {
		const uint64_t x = lb_load(base + (size_t)t * stride);
		const uint32_t ep = (uint32_t)(x >> 32);
		const uint32_t fl = (uint32_t)(x >> 30) & 3u;
		if (ep != epoch || fl == 0) {
			__builtin_amdgcn_s_sleep(1);
			continue;
		}
		excl += (uint32_t)(x & 0x3FFFFFFFu);
		if (fl == LB_INC) break;
		t--;
	}
	return excl;
}

/* Work-group exclusive scan of one value per thread (256 threads). */
__device__ __forceinline__ uint32_t wg_excl_scan(uint32_t v, uint32_t *s_tmp,
	uint32_t &total) {
	const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
	uint32_t x = v;
#pragma unroll
	for (int o = 1; o < WAVE; o <<= 1) {
		const uint32_t t = __shfl_up(x, o);
		if (lane >= o) x += t;
	}
	if (lane == WAVE - 1) s_tmp[w] = x;
	__syncthreads();
	uint32_t pre = 0, tot = 0;
#pragma unroll
	for (int i = 0; i < NWAVE; i++) {
		const uint32_t t = s_tmp[i];
		pre += (i < w) ? t : 0u;
		tot += t;
	}
	__syncthreads();
	total = tot;
	return pre + x - v;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
	for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
	return v;
}

/* Work-group sum of up to 5 counters into global atomics (one per group). */
template <int K>
__device__ __forceinline__ void wg_add(unsigned long long (&v)[K],
	unsigned long long (*s_red)[K], unsigned long long *const dst[K]) {
	const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
	for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
	if (lane == 0) {
#pragma unroll
		for (int k = 0; k < K; k++) s_red[w][k] = v[k];
	}
	__syncthreads();
	if (threadIdx.x < K) {
		unsigned long long t = 0;
		for (int i = 0; i < NWAVE; i++) t += s_red[i][threadIdx.x];
		if (t) atomicAdd(dst[threadIdx.x], t);
	}
}

__device__ __forceinline__ void set_error(Ctl *ctl, int32_t code,
	uint32_t step) {
	if (atomicCAS(&ctl->error, 0, code) == 0) ctl->err_step = step;
}

/* ------------------------------------------------------------------- init */

/* Zero the iteration counters and the grass ring; bump the look-back epoch so
 * that no word of an earlier run is taken as current. */
__global__ __launch_bounds__(WG) void k_init_reset(Ctl *ctl,
	unsigned long long *ring, uint32_t ring_n, uint32_t *ghist) {
	const uint32_t gtid = blockIdx.x * WG + threadIdx.x;
	for (uint32_t i = gtid; i < ring_n; i += gridDim.x * WG) ring[i] = 0;
	for (uint32_t i = gtid; i < MAX_PASSES * RADIX; i += gridDim.x * WG)
		ghist[i] = 0;
	if (gtid == 0) {
		ctl->n = 0;
		ctl->step = 0;
		ctl->error = 0;
		ctl->err_step = 0xFFFFFFFFu;
		ctl->live = 0;
		ctl->out_n = 0;
		ctl->epoch += 8;
		for (int i = 0; i < 8; i++) ctl->tile_ctr[i] = 0;
		ctl->dead_cells = 0;
		ctl->cd_sum = 0;
		ctl->eaten = 0;
		for (int i = 0; i < 4; i++) ctl->acc[i] = 0;
		ctl->agent_steps = 0;
		ctl->occupied = 0;
		ctl->sum_live = 0;
		ctl->sum_out = 0;
		ctl->sum_occ = 0;
		ctl->sum_eaten = 0;
	}
}

/* Cells: grass alive with probability 1/2 (draw word 0), else countdown
 * 1 + U[0, R) (word 1), Philox counter (cell, 0, INIT_CELL, 0). Follows
 * init_cell (pp_gpu.cl:368-398) and pphpc.nlogo:27-35. The grass of a cell is
 * stored as its regrow step: countdown(step s) = max(0, regrow - s). */
__global__ __launch_bounds__(WG) void k_init_cells(uint32_t *grass_at,
	unsigned long long *ring, Ctl *ctl, Params P) {
	__shared__ uint32_t s_hist[4096];
	__shared__ unsigned long long s_red[NWAVE][2];
	const bool lds_hist = P.R + 1 <= 4096;
	if (lds_hist)
		for (uint32_t i = threadIdx.x; i <= P.R; i += WG) s_hist[i] = 0;
	__syncthreads();
	unsigned long long v[2] = {0, 0};
	for (uint32_t c = blockIdx.x * WG + threadIdx.x; c < P.C;
		c += gridDim.x * WG) {
		const uint4 r = pp_philox(c, 0, PP_STREAM_INIT_CELL, 0, P.key0, P.key1);
		uint32_t g = 0;
		if (pp_bounded(r.x, 2) != 1) g = 1 + pp_bounded(r.y, P.R);
		grass_at[c] = g;
		if (g) {
			v[0] += 1;
			v[1] += g;
			if (lds_hist) atomicAdd(&s_hist[g], 1u);
			else atomicAdd(&ring[g], 1ull);
		}
	}
	unsigned long long *dst[2] = {&ctl->dead_cells, &ctl->cd_sum};
	wg_add<2>(v, s_red, dst);
	if (lds_hist) {
		__syncthreads();
		for (uint32_t i = threadIdx.x; i <= P.R; i += WG)
			if (s_hist[i]) atomicAdd(&ring[i], (unsigned long long)s_hist[i]);
	}
}

/* Agents: ids 0..init_sheep-1 are sheep, the rest wolves; x = U[0,GX) (word 0),
 * y = U[0,GY) (word 1), energy 1 + U[0, 2*gain) (word 2), Philox counter
 * (id, 0, INIT_AGENT, 0). Follows init_agent (pp_gpu.cl:406-439) and
 * pphpc.nlogo:38-53. */
__global__ __launch_bounds__(WG) void k_init_agents(uint32_t *cell,
	int32_t *energy, uint8_t *type, Ctl *ctl, Params P) {
	__shared__ unsigned long long s_red[NWAVE][4];
	unsigned long long v[4] = {0, 0, 0, 0};
	for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < P.n0;
		i += gridDim.x * WG) {
		const uint4 r = pp_philox(i, 0, PP_STREAM_INIT_AGENT, 0, P.key0, P.key1);
		const bool wolf = i >= P.init_sheep;
		const uint32_t gain = wolf ? P.wolves_gain : P.sheep_gain;
		const uint32_t x = pp_bounded(r.x, P.gx);
		const uint32_t y = pp_bounded(r.y, P.gy);
		const int32_t e = 1 + (int32_t)pp_bounded(r.z, 2 * gain);
		cell[i] = y * P.gx + x;
		energy[i] = e;
		type[i] = wolf ? PP_WOLF : PP_SHEEP;
		if (wolf) { v[1] += 1; v[3] += (unsigned long long)e; }
		else { v[0] += 1; v[2] += (unsigned long long)e; }
	}
	unsigned long long *dst[4] = {&ctl->acc[0], &ctl->acc[1], &ctl->acc[2],
		&ctl->acc[3]};
	wg_add<4>(v, s_red, dst);
}

__device__ void write_row(Ctl *ctl, pp_stats64_t *row, const Params &P) {
	row->sheep = ctl->acc[0];
	row->wolves = ctl->acc[1];
	row->grass = (unsigned long long)P.C - ctl->dead_cells;
	row->sheep_en = ctl->acc[2];
	row->wolves_en = ctl->acc[3];
	row->grass_en = ctl->cd_sum;
	row->errors = 0;
}

__global__ void k_init_done(Ctl *ctl, pp_stats64_t *stats, Params P) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	if (P.n0 > P.cap) {
		set_error(ctl, PP_OUT_OF_RESOURCES, 0);
		return;
	}
	write_row(ctl, &stats[0], P);
	for (int i = 0; i < 4; i++) ctl->acc[i] = 0;
	ctl->n = P.n0;
}

/* ------------------------------------------------------------------- move */

/* One Philox block per agent, counter (id, step, MOVE, 0): direction
 * U[0,5) from word 0 (0 stay, 1 +x, 2 -x, 3 +y, 4 -y, torus wrap: the table of
 * move_agent, pp_gpu.cl:780-803), reproduction draw U[0,100) from word 1,
 * kept as one bit (draw < probability of the agent's type) for the action
 * phase. Energy -1; the agent dies below 1 (pphpc.nlogo:66-70,
 * pp_gpu.cl:806-815). Survivors leave (key = new cell, packed value); dead
 * agents get DEAD_KEY and are dropped by the first sort pass. The histograms
 * of every radix pass are counted here, in LDS, and merged once per group. */
__global__ __launch_bounds__(WG) void k_move(const uint32_t *__restrict__ cell,
	const int32_t *__restrict__ energy, const uint8_t *__restrict__ type,
	uint32_t *__restrict__ keys, uint32_t *__restrict__ vals, Ctl *ctl,
	uint32_t *__restrict__ ghist, Params P) {
	__shared__ uint32_t s_hist[MAX_PASSES * RADIX];
	__shared__ unsigned long long s_red[NWAVE][1];
	if (ctl->error) return;
	const uint32_t n = ctl->n, step = ctl->step;
	for (uint32_t i = threadIdx.x; i < P.passes * RADIX; i += WG) s_hist[i] = 0;
	__syncthreads();
	unsigned long long live[1] = {0};
	bool overflow = false;
	const uint32_t wrap_y = P.C - P.gx;
	for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n;
		i += gridDim.x * WG) {
		const uint32_t c = cell[i];
		const int32_t e = energy[i] - 1;
		const uint32_t wolf = type[i];
		const uint4 r = pp_philox(i, step, PP_STREAM_MOVE, 0, P.key0, P.key1);
		const uint32_t dir = pp_bounded(r.x, 5);
		const uint32_t dice = pp_bounded(r.y, 100);
		const uint32_t x = c % P.gx;
		uint32_t nc = c;
		if (dir == 1) nc = (x == P.gx - 1) ? c - (P.gx - 1) : c + 1;
		else if (dir == 2) nc = (x == 0) ? c + (P.gx - 1) : c - 1;
		else if (dir == 3) nc = (c >= wrap_y) ? c - wrap_y : c + P.gx;
		else if (dir == 4) nc = (c < P.gx) ? c + wrap_y : c - P.gx;
		uint32_t k = DEAD_KEY, v = 0;
		if (e >= 1) {
			const uint32_t prob = wolf ? P.wolves_prob : P.sheep_prob;
			overflow |= e > (int32_t)V_ENERGY;
			k = nc;
			v = ((uint32_t)e & V_ENERGY) | (wolf ? V_WOLF : 0u)
				| (dice < prob ? V_REPRO : 0u);
			live[0] += 1;
			for (uint32_t p = 0; p < P.passes; p++)
				atomicAdd(&s_hist[p * RADIX + ((nc >> (8 * p)) & 0xFFu)], 1u);
		}
		keys[i] = k;
		vals[i] = v;
	}
	if (overflow) set_error(ctl, PP_OUT_OF_RESOURCES, step);
	/* live count: one atomic per group */
	{
		const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
		const unsigned long long t = wave_sum(live[0]);
		if (lane == 0) s_red[w][0] = t;
		__syncthreads();
		if (threadIdx.x == 0) {
			unsigned long long s = 0;
			for (int i = 0; i < NWAVE; i++) s += s_red[i][0];
			if (s) atomicAdd(&ctl->live, (uint32_t)s);
		}
	}
	for (uint32_t i = threadIdx.x; i < P.passes * RADIX; i += WG) {
		const uint32_t h = s_hist[i];
		if (h) atomicAdd(&ghist[i], h);
	}
}

/* ------------------------------------------------------------------- sort */

/* One LSD pass of a stable radix sort of (key, value) pairs on 8 key bits at
 * `shift`, onesweep style: work-groups claim tiles in order from an atomic
 * counter; each ranks its tile's keys in LDS with wave64 ballot matching
 * (stable: input order within the tile), publishes per-digit tile counts and
 * obtains the counts of all earlier tiles by decoupled look-back, then writes
 * the tile out digit-contiguously through LDS. The first pass also drops dead
 * agents (DEAD_KEY), so later passes see only survivors. Sorting survivors
 * stably by cell is sorting by the key (cell << 32 | agent id). */
template <int KPT>
__global__ __launch_bounds__(WG) void k_sort_pass(
	const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
	uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, Ctl *ctl,
	const uint32_t *__restrict__ ghist, unsigned long long *lbuf,
	uint32_t pass) {
	constexpr uint32_t TILE = WG * KPT;
	__shared__ uint32_t s_keys[TILE];
	__shared__ uint32_t s_vals[TILE];
	__shared__ uint32_t s_wh[NWAVE][RADIX + 1];
	__shared__ uint32_t s_bstart[RADIX];
	__shared__ uint32_t s_tstart[RADIX];
	__shared__ int32_t s_dst[RADIX];
	__shared__ uint32_t s_tmp[NWAVE];
	__shared__ uint32_t s_tile;

	if (ctl->error) return;
	const uint32_t n = pass == 0 ? ctl->n : ctl->live;
	const uint32_t ntiles = (n + TILE - 1) / TILE;
	const uint32_t epoch = ctl->epoch + pass + 1;
	const uint32_t shift = 8 * pass;
	const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
	const uint64_t lt_mask = (1ull << lane) - 1ull;

	{
		uint32_t tot;
		s_bstart[tid] = wg_excl_scan(ghist[pass * RADIX + tid], s_tmp, tot);
	}

	for (;;) {
		if (tid == 0) s_tile = atomicAdd(&ctl->tile_ctr[pass], 1u);
		for (int i = tid; i < NWAVE * (RADIX + 1); i += WG) (&s_wh[0][0])[i] = 0;
		__syncthreads();
		const uint32_t tile = s_tile;
		if (tile >= ntiles) break;

		/* Load: wave w owns [tile*TILE + w*64*KPT, +64*KPT), element j of lane
		 * l at offset j*64 + l (coalesced). */
		const uint32_t base = tile * TILE + w * (WAVE * KPT);
		uint32_t key[KPT], val[KPT], dig[KPT], loc[KPT];
#pragma unroll
		for (int j = 0; j < KPT; j++) {
			const uint32_t idx = base + j * WAVE + lane;
			uint32_t k = DEAD_KEY, v = 0;
			if (idx < n) {
				k = kin[idx];
				v = vin[idx];
			}
			key[j] = k;
			val[j] = v;
			dig[j] = (k == DEAD_KEY) ? (uint32_t)RADIX : ((k >> shift) & 0xFFu);
		}
		/* Rank within the wave: lanes with an equal 9-bit digit (digit 256 =
		 * dead or padding) found by 9 ballots; order (j, lane) = input order. */
#pragma unroll
		for (int j = 0; j < KPT; j++) {
			const uint32_t d = dig[j];
			uint64_t peers = ~0ull;
#pragma unroll
			for (int b = 0; b < 9; b++) {
				const bool bit = (d >> b) & 1u;
				const uint64_t bal = __ballot(bit);
				peers &= bit ? bal : ~bal;
			}
			const uint32_t rank = __popcll(peers & lt_mask);
			const uint32_t cnt = __popcll(peers);
			const uint32_t prev = s_wh[w][d];
			if (rank == 0) s_wh[w][d] = prev + cnt;
			loc[j] = prev + rank;
		}
		__syncthreads();

		/* Digit tid: per-wave exclusive offsets and the tile count. */
		const uint32_t c0 = s_wh[0][tid], c1 = s_wh[1][tid];
		const uint32_t c2 = s_wh[2][tid], c3 = s_wh[3][tid];
		s_wh[0][tid] = 0;
		s_wh[1][tid] = c0;
		s_wh[2][tid] = c0 + c1;
		s_wh[3][tid] = c0 + c1 + c2;
		const uint32_t cnt_d = c0 + c1 + c2 + c3;
		uint32_t tile_live;
		const uint32_t tstart = wg_excl_scan(cnt_d, s_tmp, tile_live);
		s_tstart[tid] = tstart;

		unsigned long long *my = lbuf + (size_t)tile * RADIX + tid;
		uint32_t excl = 0;
		if (tile == 0) {
			lb_store(my, lb_pack(epoch, LB_INC, cnt_d));
		} else {
			lb_store(my, lb_pack(epoch, LB_AGG, cnt_d));
			excl = lookback(lbuf + tid, RADIX, tile, epoch);
			lb_store(my, lb_pack(epoch, LB_INC, excl + cnt_d));
		}
		s_dst[tid] = (int32_t)(s_bstart[tid] + excl) - (int32_t)tstart;
		__syncthreads();

		/* Stage the tile in digit order in LDS. */
#pragma unroll
		for (int j = 0; j < KPT; j++) {
			const uint32_t d = dig[j];
			if (d < RADIX) {
				const uint32_t q = s_tstart[d] + s_wh[w][d] + loc[j];
				s_keys[q] = key[j];
				s_vals[q] = val[j];
			}
		}
		__syncthreads();
		for (uint32_t q = tid; q < tile_live; q += WG) {
			const uint32_t k = s_keys[q];
			const uint32_t d = (k >> shift) & 0xFFu;
			const uint32_t dst = (uint32_t)(s_dst[d] + (int32_t)q);
			kout[dst] = k;
			vout[dst] = s_vals[q];
		}
		__syncthreads();
	}
}

/* ----------------------------------------------------------------- action */

/* Agent actions for the agents sorted by cell (ascending agent id within a
 * cell). Thread p handles sorted position p; when p starts a run of equal
 * cells (the segment that find_cell_idx, pp_gpu.cl:834-867, delimits), the
 * thread acts for the whole cell in sequence:
 *   1. Fisher-Yates permutation of the segment, the draw for position i being
 *      word m%4 of Philox block (cell, step, SHUFFLE, m/4), m = k-1-i (random
 *      acting order: Cell.java:246, pp_cpu.cl:236-259);
 *   2. in that order, each agent not yet eaten: a sheep eats the grass if the
 *      grass is alive (countdown 0 after this iteration's regrowth) and gains
 *      SHEEP_GAIN (Sheep.java:66-83); a wolf eats the first sheep of the
 *      permuted segment not yet eaten, whether or not it has acted, and gains
 *      WOLVES_GAIN (Wolf.java:66-92); then, if energy > threshold and the
 *      reproduction draw passed, the energy is split, the child getting
 *      floor(e/2) (AbstractAgent.java:139-167). Newborns are not in the
 *      segment, so they neither act nor get eaten (pp_cpu.cl:841-845).
 * The next state is written compacted: cells ascending, within a cell the
 * surviving agents in acting order each followed by its child. Output offsets
 * come from a work-group scan of the per-cell output counts plus decoupled
 * look-back across tiles. Counts and energy sums of the new state are reduced
 * here; grass eaten is counted for the ring update in k_finalize. */
__global__ __launch_bounds__(WG) void k_action(
	const uint32_t *__restrict__ keys, uint32_t *vals,
	uint32_t *__restrict__ ocell, int32_t *__restrict__ oenergy,
	uint8_t *__restrict__ otype, uint32_t *grass_at, Ctl *ctl,
	unsigned long long *lbuf, Params P) {
	__shared__ uint32_t s_tmp[NWAVE];
	__shared__ uint32_t s_tile;
	__shared__ uint32_t s_off;
	__shared__ unsigned long long s_red[NWAVE][6];

	if (ctl->error) return;
	const uint32_t L = ctl->live, step = ctl->step;
	const uint32_t ntiles = (L + WG - 1) / WG;
	const uint32_t epoch = ctl->epoch + ACTION_SLOT + 1;
	const uint32_t alive_at = step + 1;
	const int tid = threadIdx.x;
	unsigned long long acc[6] = {0, 0, 0, 0, 0, 0};
	bool overflow = false;

	for (;;) {
		if (tid == 0) s_tile = atomicAdd(&ctl->tile_ctr[ACTION_SLOT], 1u);
		__syncthreads();
		const uint32_t tile = s_tile;
		if (tile >= ntiles) break;
		const uint32_t p = tile * WG + tid;
		uint32_t c = 0, end = 0, nout = 0;
		bool head = false;
		if (p < L) {
			c = keys[p];
			head = (p == 0) || (keys[p - 1] != c);
		}
		if (head) {
			end = p + 1;
			while (end < L && keys[end] == c) end++;
			const uint32_t k = end - p;
			acc[5] += 1;
			if (k > 1) {
				uint4 r = make_uint4(0, 0, 0, 0);
				for (uint32_t i = k - 1, m = 0; i > 0; i--, m++) {
					if ((m & 3u) == 0)
						r = pp_philox(c, step, PP_STREAM_SHUFFLE, m >> 2, P.key0,
							P.key1);
					const uint32_t word = (m & 3u) == 0 ? r.x
						: (m & 3u) == 1 ? r.y : (m & 3u) == 2 ? r.z : r.w;
					const uint32_t j = pp_bounded(word, i + 1);
					if (j != i) {
						const uint32_t a = vals[p + i];
						vals[p + i] = vals[p + j];
						vals[p + j] = a;
					}
				}
			}
			bool grass = grass_at[c] <= alive_at;
			bool ate = false;
			for (uint32_t q = p; q < end; q++) {
				const uint32_t v = vals[q];
				if (v & V_KILLED) continue;
				int32_t e = (int32_t)(v & V_ENERGY);
				uint32_t thr;
				if (!(v & V_WOLF)) {
					if (grass) {
						grass = false;
						ate = true;
						e += (int32_t)P.sheep_gain;
					}
					thr = P.sheep_thr;
				} else {
					for (uint32_t o = p; o < end; o++) {
						const uint32_t u = vals[o];
						if (!(u & (V_WOLF | V_KILLED))) {
							vals[o] = u | V_KILLED;
							e += (int32_t)P.wolves_gain;
							break;
						}
					}
					thr = P.wolves_thr;
				}
				uint32_t nv = v & V_WOLF;
				if (e > (int32_t)thr && (v & V_REPRO)) {
					nv |= V_CHILD | ((e & 1) ? V_ODD : 0u);
					e -= e / 2;
				}
				overflow |= e > (int32_t)V_ENERGY;
				vals[q] = nv | ((uint32_t)e & V_ENERGY);
			}
			if (ate) {
				grass_at[c] = step + 1 + P.R;
				acc[4] += 1;
			}
			for (uint32_t q = p; q < end; q++) {
				const uint32_t v = vals[q];
				nout += ((v & V_KILLED) == 0) + ((v & V_CHILD) != 0);
			}
		}

		uint32_t tile_total;
		const uint32_t local = wg_excl_scan(nout, s_tmp, tile_total);
		if (tid == 0) {
			unsigned long long *my = lbuf + tile;
			uint32_t excl = 0;
			if (tile == 0) {
				lb_store(my, lb_pack(epoch, LB_INC, tile_total));
			} else {
				lb_store(my, lb_pack(epoch, LB_AGG, tile_total));
				excl = lookback(lbuf, 1, tile, epoch);
				lb_store(my, lb_pack(epoch, LB_INC, excl + tile_total));
			}
			s_off = excl;
			if (tile == ntiles - 1) ctl->out_n = excl + tile_total;
		}
		__syncthreads();
		if (head) {
			uint32_t o = s_off + local;
			for (uint32_t q = p; q < end; q++) {
				const uint32_t v = vals[q];
				const int32_t e = (int32_t)(v & V_ENERGY);
				const uint8_t t = (v & V_WOLF) ? PP_WOLF : PP_SHEEP;
				if (!(v & V_KILLED)) {
					if (o < P.cap) {
						ocell[o] = c;
						oenergy[o] = e;
						otype[o] = t;
					}
					o++;
					acc[t] += 1;
					acc[2 + t] += (unsigned long long)e;
				}
				if (v & V_CHILD) {
					const int32_t ce = (v & V_ODD) ? e - 1 : e;
					if (o < P.cap) {
						ocell[o] = c;
						oenergy[o] = ce;
						otype[o] = t;
					}
					o++;
					acc[t] += 1;
					acc[2 + t] += (unsigned long long)ce;
				}
			}
		}
		__syncthreads();
	}
	if (overflow) set_error(ctl, PP_OUT_OF_RESOURCES, step);
	unsigned long long *dst[6] = {&ctl->acc[0], &ctl->acc[1], &ctl->acc[2],
		&ctl->acc[3], &ctl->eaten, &ctl->occupied};
	wg_add<6>(acc, s_red, dst);
}

/* --------------------------------------------------------------- finalize */

/* End of an iteration s -> s+1 (one work-group). Grass statistics come from
 * the regrow-step encoding without touching the grid: with D dead cells and
 * countdown sum S at state s, every dead countdown drops by one (S -= D), the
 * cells whose regrow step is s+1 come alive (ring slot (s+1) mod (R+1)), and
 * the E cells eaten this iteration restart at R (slot s mod (R+1)): exactly
 * the grass kernel (pp_gpu.cl:448-476) plus the sums of reduce_grass1/2
 * (pp_gpu.cl:485-583). Then the statistics row s+1 is written, the capacity
 * checked (pp_gpu.c:847-852) and the per-iteration counters cleared. */
__global__ __launch_bounds__(WG) void k_finalize(Ctl *ctl,
	unsigned long long *ring, uint32_t *ghist, pp_stats64_t *stats,
	Params P) {
	if (threadIdx.x == 0) {
		if (ctl->error == 0) {
			const uint32_t s = ctl->step;
			if (ctl->out_n > P.cap) {
				set_error(ctl, PP_OUT_OF_RESOURCES, s);
			} else {
				const uint32_t M = P.R + 1;
				unsigned long long D = ctl->dead_cells, S = ctl->cd_sum;
				S -= D;
				const uint32_t expire = (s + 1) % M;
				D -= ring[expire];
				ring[expire] = 0;
				const unsigned long long E = ctl->eaten;
				ring[s % M] += E;
				D += E;
				S += E * (unsigned long long)P.R;
				ctl->dead_cells = D;
				ctl->cd_sum = S;
				write_row(ctl, &stats[(s + 1) % STATS_RING], P);
				ctl->agent_steps += ctl->n;
				ctl->sum_live += ctl->live;
				ctl->sum_out += ctl->out_n;
				ctl->sum_occ += ctl->occupied;
				ctl->sum_eaten += ctl->eaten;
				ctl->n = ctl->out_n;
				ctl->step = s + 1;
			}
			for (int i = 0; i < 4; i++) ctl->acc[i] = 0;
			ctl->eaten = 0;
			ctl->occupied = 0;
			ctl->live = 0;
			ctl->out_n = 0;
			ctl->epoch += 8;
			for (int i = 0; i < 8; i++) ctl->tile_ctr[i] = 0;
		}
	}
	for (uint32_t i = threadIdx.x; i < MAX_PASSES * RADIX; i += WG) ghist[i] = 0;
}

/* ------------------------------------------------------------ host layer */

enum KernelId {
	KI_INIT_RESET,
	KI_INIT_CELLS,
	KI_INIT_AGENTS,
	KI_INIT_DONE,
	KI_MOVE,
	KI_SORT_PASS,
	KI_ACTION,
	KI_FINALIZE,
	KI_COUNT
};

const char *const kernel_names[KI_COUNT] = {
	"init_reset", "init_cells", "init_agents", "init_done",
	"move", "sort_pass", "action", "finalize",
};

struct PendingEvent {
	int kernel;
	hipEvent_t start, stop;
};

} /* namespace */

struct pp_sim {
	pp_params_t prm;
	Params P;
	uint64_t seed;
	uint32_t cap;
	int device;
	int num_cu;
	hipStream_t stream;

	uint32_t *cell;
	int32_t *energy;
	uint8_t *type;
	uint32_t *keys[2];
	uint32_t *vals[2];
	uint32_t *grass_at;
	unsigned long long *ring;
	uint32_t ring_n;
	uint32_t *ghist;
	unsigned long long *lb_sort;
	unsigned long long *lb_act;
	Ctl *ctl;
	pp_stats64_t *stats_dev;
	pp_stats64_t *stats_host; /* pinned, STATS_RING rows */
	Ctl *ctl_host;            /* pinned mirror */

	int kpt;
	uint32_t sort_tile;
	uint32_t grid_move, grid_sort, grid_act, grid_cells, grid_agents;

	hipGraphExec_t graph_exec;
	int graph_steps;

	int profiling;
	std::vector<PendingEvent> pending;
	std::vector<hipEvent_t> free_events;
	uint64_t prof_launches[KI_COUNT];
	double prof_ns[KI_COUNT];

	int initialised;
};

#define HIP_OK(expr)                                                        \
	do {                                                                    \
		hipError_t e_ = (expr);                                             \
		if (e_ != hipSuccess) {                                             \
			fprintf(stderr, "pphpc: HIP error %s at %s:%d: %s\n",           \
				hipGetErrorName(e_), __FILE__, __LINE__, #expr);            \
			return PP_LIBRARY_ERROR;                                        \
		}                                                                   \
	} while (0)

static hipEvent_t event_get(pp_sim_t *s) {
	if (!s->free_events.empty()) {
		hipEvent_t e = s->free_events.back();
		s->free_events.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	if (hipEventCreate(&e) != hipSuccess) return nullptr;
	return e;
}

/* Profiling wrappers: bracket a launch with events on the sim's stream. */
static void prof_begin(pp_sim_t *s, int kernel, PendingEvent &pe) {
	pe.kernel = kernel;
	pe.start = pe.stop = nullptr;
	if (!s->profiling) return;
	pe.start = event_get(s);
	pe.stop = event_get(s);
	if (pe.start) (void)hipEventRecord(pe.start, s->stream);
}

static void prof_end(pp_sim_t *s, PendingEvent &pe) {
	if (!s->profiling || !pe.start || !pe.stop) return;
	(void)hipEventRecord(pe.stop, s->stream);
	s->pending.push_back(pe);
}

static int prof_collect(pp_sim_t *s) {
	for (auto &pe : s->pending) {
		float ms = 0.f;
		HIP_OK(hipEventSynchronize(pe.stop));
		HIP_OK(hipEventElapsedTime(&ms, pe.start, pe.stop));
		s->prof_ns[pe.kernel] += 1e6 * (double)ms;
		s->prof_launches[pe.kernel] += 1;
		s->free_events.push_back(pe.start);
		s->free_events.push_back(pe.stop);
	}
	s->pending.clear();
	return PP_SUCCESS;
}

template <int KPT>
static void launch_sort_pass(pp_sim_t *s, uint32_t pass) {
	const int src = pass & 1, dst = src ^ 1;
	k_sort_pass<KPT><<<s->grid_sort, WG, 0, s->stream>>>(s->keys[src],
		s->vals[src], s->keys[dst], s->vals[dst], s->ctl, s->ghist, s->lb_sort,
		pass);
}

/* Enqueue one iteration (no synchronisation). */
static int enqueue_step(pp_sim_t *s) {
	PendingEvent pe;
	prof_begin(s, KI_MOVE, pe);
	k_move<<<s->grid_move, WG, 0, s->stream>>>(s->cell, s->energy, s->type,
		s->keys[0], s->vals[0], s->ctl, s->ghist, s->P);
	prof_end(s, pe);
	for (uint32_t pass = 0; pass < s->P.passes; pass++) {
		prof_begin(s, KI_SORT_PASS, pe);
		switch (s->kpt) {
		case 4: launch_sort_pass<4>(s, pass); break;
		case 16: launch_sort_pass<16>(s, pass); break;
		default: launch_sort_pass<8>(s, pass); break;
		}
		prof_end(s, pe);
	}
	const int sorted = s->P.passes & 1;
	prof_begin(s, KI_ACTION, pe);
	k_action<<<s->grid_act, WG, 0, s->stream>>>(s->keys[sorted],
		s->vals[sorted], s->cell, s->energy, s->type, s->grass_at, s->ctl,
		s->lb_act, s->P);
	prof_end(s, pe);
	prof_begin(s, KI_FINALIZE, pe);
	k_finalize<<<1, WG, 0, s->stream>>>(s->ctl, s->ring, s->ghist,
		s->stats_dev, s->P);
	prof_end(s, pe);
	HIP_OK(hipGetLastError());
	return PP_SUCCESS;
}

static int read_ctl(pp_sim_t *s) {
	HIP_OK(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(Ctl),
		hipMemcpyDeviceToHost, s->stream));
	HIP_OK(hipStreamSynchronize(s->stream));
	return prof_collect(s);
}

static uint32_t ceil_div(uint64_t a, uint64_t b) {
	return (uint32_t)((a + b - 1) / b);
}

static uint32_t min_u32(uint32_t a, uint32_t b) { return a < b ? a : b; }

static void sim_free(pp_sim_t *s) {
	if (!s) return;
	if (s->device >= 0) (void)hipSetDevice(s->device);
	if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
	for (auto &pe : s->pending) {
		(void)hipEventDestroy(pe.start);
		(void)hipEventDestroy(pe.stop);
	}
	for (auto e : s->free_events) (void)hipEventDestroy(e);
	(void)hipFree(s->cell);
	(void)hipFree(s->energy);
	(void)hipFree(s->type);
	for (int i = 0; i < 2; i++) {
		(void)hipFree(s->keys[i]);
		(void)hipFree(s->vals[i]);
	}
	(void)hipFree(s->grass_at);
	(void)hipFree(s->ring);
	(void)hipFree(s->ghist);
	(void)hipFree(s->lb_sort);
	(void)hipFree(s->lb_act);
	(void)hipFree(s->ctl);
	(void)hipFree(s->stats_dev);
	if (s->stats_host) (void)hipHostFree(s->stats_host);
	if (s->ctl_host) (void)hipHostFree(s->ctl_host);
	if (s->stream) (void)hipStreamDestroy(s->stream);
	delete s;
}

static int sim_alloc(pp_sim_t *s) {
	const uint64_t A = s->cap, C = s->P.C;
	HIP_OK(hipMalloc(&s->cell, A * 4));
	HIP_OK(hipMalloc(&s->energy, A * 4));
	HIP_OK(hipMalloc(&s->type, A));
	for (int i = 0; i < 2; i++) {
		HIP_OK(hipMalloc(&s->keys[i], A * 4));
		HIP_OK(hipMalloc(&s->vals[i], A * 4));
	}
	HIP_OK(hipMalloc(&s->grass_at, C * 4));
	s->ring_n = s->P.R + 1;
	HIP_OK(hipMalloc(&s->ring, (size_t)s->ring_n * 8));
	HIP_OK(hipMalloc(&s->ghist, MAX_PASSES * RADIX * 4));
	const uint64_t sort_tiles = ceil_div(A, s->sort_tile);
	const uint64_t act_tiles = ceil_div(A, WG);
	HIP_OK(hipMalloc(&s->lb_sort, sort_tiles * RADIX * 8));
	HIP_OK(hipMalloc(&s->lb_act, act_tiles * 8));
	HIP_OK(hipMalloc(&s->ctl, sizeof(Ctl)));
	HIP_OK(hipMalloc(&s->stats_dev, STATS_RING * sizeof(pp_stats64_t)));
	HIP_OK(hipHostMalloc(&s->stats_host, STATS_RING * sizeof(pp_stats64_t),
		hipHostMallocDefault));
	HIP_OK(hipHostMalloc(&s->ctl_host, sizeof(Ctl), hipHostMallocDefault));
	/* Zeroed once at creation, outside any graph: epoch 0 never matches. */
	HIP_OK(hipMemset(s->lb_sort, 0, sort_tiles * RADIX * 8));
	HIP_OK(hipMemset(s->lb_act, 0, act_tiles * 8));
	HIP_OK(hipMemset(s->ctl, 0, sizeof(Ctl)));
	HIP_OK(hipMemset(s->ghist, 0, MAX_PASSES * RADIX * 4));
	HIP_OK(hipDeviceSynchronize());
	return PP_SUCCESS;
}

extern "C" {

int pp_sim_create(const pp_params_t *p, uint64_t seed, uint64_t capacity,
	int device, pp_sim_t **out) {
	if (!p || !out) return PP_INVALID_ARGS;
	*out = nullptr;
	const uint64_t C = (uint64_t)p->grid_x * p->grid_y;
	const uint64_t n0 = (uint64_t)p->init_sheep + p->init_wolves;
	if (p->grid_x == 0 || p->grid_y == 0 || C > 0xFFFFFFFEull)
		return PP_INVALID_ARGS;
	if (p->grass_restart == 0 || p->grass_restart > (1u << 24))
		return PP_INVALID_ARGS;
	if (p->sheep_gain_from_food > (1u << 20) || p->wolves_gain_from_food > (1u << 20))
		return PP_INVALID_ARGS;
	if (capacity == 0) {
		capacity = 4 * n0;
		if (capacity < PP_DEFAULT_MAX_AGENTS) capacity = PP_DEFAULT_MAX_AGENTS;
	}
	/* Look-back words carry 30-bit counts; an iteration can at most double
	 * the survivors, so the capacity stays below 2^29. */
	if (capacity < n0 || capacity >= (1ull << 29)) return PP_INVALID_ARGS;

	pp_sim_t *s = new pp_sim();
	memset(s->prof_launches, 0, sizeof(s->prof_launches));
	memset(s->prof_ns, 0, sizeof(s->prof_ns));
	s->prm = *p;
	s->prm.grid_xy = (uint32_t)C;
	s->seed = seed;
	s->cap = (uint32_t)capacity;
	s->device = device;
	if (device < 0) {
		if (hipGetDevice(&s->device) != hipSuccess) {
			delete s;
			return PP_LIBRARY_ERROR;
		}
	}
	if (hipSetDevice(s->device) != hipSuccess) {
		s->device = -1;
		sim_free(s);
		return PP_LIBRARY_ERROR;
	}
	if (hipDeviceGetAttribute(&s->num_cu, hipDeviceAttributeMultiprocessorCount,
			s->device) != hipSuccess || s->num_cu <= 0)
		s->num_cu = 256;
	if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
		sim_free(s);
		return PP_LIBRARY_ERROR;
	}

	Params &P = s->P;
	P.key0 = (uint32_t)seed;
	P.key1 = (uint32_t)(seed >> 32);
	P.gx = p->grid_x;
	P.gy = p->grid_y;
	P.C = (uint32_t)C;
	P.R = p->grass_restart;
	P.sheep_gain = p->sheep_gain_from_food;
	P.wolves_gain = p->wolves_gain_from_food;
	P.sheep_thr = p->sheep_reproduce_threshold;
	P.wolves_thr = p->wolves_reproduce_threshold;
	P.sheep_prob = p->sheep_reproduce_prob;
	P.wolves_prob = p->wolves_reproduce_prob;
	P.init_sheep = p->init_sheep;
	P.n0 = (uint32_t)n0;
	P.cap = s->cap;
	uint32_t bits = 0;
	while (bits < 32 && ((C - 1) >> bits) != 0) bits++;
	P.passes = bits == 0 ? 1 : (bits + 7) / 8;

	s->kpt = 8;
	const char *kpt_env = getenv("PPHPC_SORT_KPT");
	if (kpt_env) {
		int v = atoi(kpt_env);
		if (v == 4 || v == 8 || v == 16) s->kpt = v;
	}
	s->sort_tile = WG * s->kpt;
	/* Persistent grids: enough groups to fill every CU several times, never
	 * more than the tiles of a full-capacity population. */
	s->grid_move = min_u32(ceil_div(s->cap, WG), 8 * s->num_cu);
	s->grid_sort = min_u32(ceil_div(s->cap, s->sort_tile), 4 * s->num_cu);
	s->grid_act = min_u32(ceil_div(s->cap, WG), 8 * s->num_cu);
	s->grid_cells = min_u32(ceil_div(C, WG), 8 * s->num_cu);
	s->grid_agents = min_u32(ceil_div(n0 ? n0 : 1, WG), 8 * s->num_cu);
	if (s->grid_move == 0) s->grid_move = 1;
	if (s->grid_sort == 0) s->grid_sort = 1;
	if (s->grid_act == 0) s->grid_act = 1;

	int rc = sim_alloc(s);
	if (rc != PP_SUCCESS) {
		sim_free(s);
		return rc;
	}
	*out = s;
	return PP_SUCCESS;
}

void pp_sim_destroy(pp_sim_t *s) { sim_free(s); }

int pp_sim_init(pp_sim_t *s, pp_stats64_t *stats0) {
	if (!s) return PP_INVALID_ARGS;
	HIP_OK(hipSetDevice(s->device));
	PendingEvent pe;
	prof_begin(s, KI_INIT_RESET, pe);
	k_init_reset<<<min_u32(ceil_div(s->ring_n, WG), 1024), WG, 0, s->stream>>>(
		s->ctl, s->ring, s->ring_n, s->ghist);
	prof_end(s, pe);
	prof_begin(s, KI_INIT_CELLS, pe);
	k_init_cells<<<s->grid_cells, WG, 0, s->stream>>>(s->grass_at, s->ring,
		s->ctl, s->P);
	prof_end(s, pe);
	prof_begin(s, KI_INIT_AGENTS, pe);
	k_init_agents<<<s->grid_agents, WG, 0, s->stream>>>(s->cell, s->energy,
		s->type, s->ctl, s->P);
	prof_end(s, pe);
	prof_begin(s, KI_INIT_DONE, pe);
	k_init_done<<<1, 64, 0, s->stream>>>(s->ctl, s->stats_dev, s->P);
	prof_end(s, pe);
	HIP_OK(hipGetLastError());
	HIP_OK(hipMemcpyAsync(s->stats_host, s->stats_dev, sizeof(pp_stats64_t),
		hipMemcpyDeviceToHost, s->stream));
	int rc = read_ctl(s);
	if (rc) return rc;
	if (s->ctl_host->error) return s->ctl_host->error;
	if (stats0) *stats0 = s->stats_host[0];
	s->initialised = 1;
	return PP_SUCCESS;
}

/* Capture `steps` iterations into one executable graph (kernels only; no
 * memset or copy nodes). */
static int build_graph(pp_sim_t *s, int steps) {
	hipGraph_t g = nullptr;
	HIP_OK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
	int rc = PP_SUCCESS;
	for (int i = 0; i < steps && rc == PP_SUCCESS; i++) rc = enqueue_step(s);
	hipError_t e = hipStreamEndCapture(s->stream, &g);
	if (rc != PP_SUCCESS) {
		if (g) (void)hipGraphDestroy(g);
		return rc;
	}
	HIP_OK(e);
	e = hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0);
	(void)hipGraphDestroy(g);
	HIP_OK(e);
	s->graph_steps = steps;
	return PP_SUCCESS;
}

static int copy_rows(pp_sim_t *s, uint32_t first_row, uint32_t nrows) {
	/* rows first_row .. first_row + nrows - 1 of the ring (may wrap) */
	uint32_t a = first_row % STATS_RING;
	uint32_t n1 = min_u32(nrows, STATS_RING - a);
	HIP_OK(hipMemcpyAsync(s->stats_host + a, s->stats_dev + a,
		(size_t)n1 * sizeof(pp_stats64_t), hipMemcpyDeviceToHost, s->stream));
	if (nrows > n1)
		HIP_OK(hipMemcpyAsync(s->stats_host, s->stats_dev,
			(size_t)(nrows - n1) * sizeof(pp_stats64_t), hipMemcpyDeviceToHost,
			s->stream));
	return PP_SUCCESS;
}

int pp_sim_run(pp_sim_t *s, uint32_t iters, pp_stats64_t *out) {
	if (!s || !s->initialised) return PP_INVALID_ARGS;
	HIP_OK(hipSetDevice(s->device));
	if (s->ctl_host->error) return s->ctl_host->error;
	const uint64_t last = (uint64_t)s->ctl_host->step + iters + s->P.R + 1;
	if (last >= 0xFFFFFFFFull) return PP_INVALID_ARGS;
	const int use_graph = !s->profiling && getenv("PPHPC_NO_GRAPH") == nullptr;
	uint32_t done = 0;
	while (done < iters) {
		const uint32_t chunk = min_u32(iters - done, STATS_RING / 2);
		const uint32_t step0 = s->ctl_host->step;
		uint32_t i = 0;
		if (use_graph && chunk >= 8) {
			if (!s->graph_exec) {
				int rc = build_graph(s, 8);
				if (rc) return rc;
			}
			for (; i + (uint32_t)s->graph_steps <= chunk; i += s->graph_steps)
				HIP_OK(hipGraphLaunch(s->graph_exec, s->stream));
		}
		for (; i < chunk; i++) {
			int rc = enqueue_step(s);
			if (rc) return rc;
		}
		int rc = copy_rows(s, step0 + 1, chunk);
		if (rc) return rc;
		rc = read_ctl(s);
		if (rc) return rc;
		const uint32_t got = s->ctl_host->step - step0;
		if (out)
			for (uint32_t r = 0; r < got; r++)
				out[done + r] = s->stats_host[(step0 + 1 + r) % STATS_RING];
		done += got;
		if (s->ctl_host->error) return s->ctl_host->error;
		if (got != chunk) return PP_LIBRARY_ERROR;
	}
	return PP_SUCCESS;
}

int pp_sim_step(pp_sim_t *s, pp_stats64_t *out) {
	return pp_sim_run(s, 1, out);
}

int pp_sim_simulate(pp_sim_t *s, pp_stats64_t *out) {
	if (!s) return PP_INVALID_ARGS;
	int rc = pp_sim_init(s, out);
	if (rc) return rc;
	return pp_sim_run(s, s->prm.iters, out ? out + 1 : nullptr);
}

uint32_t pp_sim_step_count(const pp_sim_t *s) {
	return s ? s->ctl_host->step : 0;
}

uint64_t pp_sim_num_agents(const pp_sim_t *s) {
	return s ? s->ctl_host->n : 0;
}

uint32_t pp_sim_error_step(const pp_sim_t *s) {
	return (s && s->ctl_host->error) ? s->ctl_host->err_step : 0xFFFFFFFFu;
}

uint64_t pp_sim_agent_steps(const pp_sim_t *s) {
	return s ? s->ctl_host->agent_steps : 0;
}

uint64_t pp_sim_capacity(const pp_sim_t *s) { return s ? s->cap : 0; }

int pp_sim_counters(const pp_sim_t *s, uint64_t out[5]) {
	if (!s || !out) return PP_INVALID_ARGS;
	out[0] = s->ctl_host->agent_steps;
	out[1] = s->ctl_host->sum_live;
	out[2] = s->ctl_host->sum_out;
	out[3] = s->ctl_host->sum_occ;
	out[4] = s->ctl_host->sum_eaten;
	return PP_SUCCESS;
}

int pp_sim_get_state(pp_sim_t *s, uint32_t *cell, int32_t *energy,
	uint8_t *type, uint32_t *grass) {
	if (!s || !s->initialised) return PP_INVALID_ARGS;
	HIP_OK(hipSetDevice(s->device));
	int rc = read_ctl(s);
	if (rc) return rc;
	const size_t n = s->ctl_host->n;
	if (cell && n) HIP_OK(hipMemcpy(cell, s->cell, n * 4, hipMemcpyDeviceToHost));
	if (energy && n)
		HIP_OK(hipMemcpy(energy, s->energy, n * 4, hipMemcpyDeviceToHost));
	if (type && n) HIP_OK(hipMemcpy(type, s->type, n, hipMemcpyDeviceToHost));
	if (grass) {
		HIP_OK(hipMemcpy(grass, s->grass_at, (size_t)s->P.C * 4,
			hipMemcpyDeviceToHost));
		const uint32_t st = s->ctl_host->step;
		for (uint32_t c = 0; c < s->P.C; c++)
			grass[c] = grass[c] > st ? grass[c] - st : 0;
	}
	return PP_SUCCESS;
}

int pp_sim_set_profiling(pp_sim_t *s, int enable) {
	if (!s) return PP_INVALID_ARGS;
	s->profiling = enable ? 1 : 0;
	return PP_SUCCESS;
}

int pp_sim_prof_count(const pp_sim_t *s) { return s ? KI_COUNT : 0; }

int pp_sim_prof_get(pp_sim_t *s, int i, const char **name, uint64_t *launches,
	double *total_ns) {
	if (!s || i < 0 || i >= KI_COUNT) return PP_INVALID_ARGS;
	int rc = prof_collect(s);
	if (rc) return rc;
	if (name) *name = kernel_names[i];
	if (launches) *launches = s->prof_launches[i];
	if (total_ns) *total_ns = s->prof_ns[i];
	return PP_SUCCESS;
}

int pp_sim_prof_reset(pp_sim_t *s) {
	if (!s) return PP_INVALID_ARGS;
	int rc = prof_collect(s);
	memset(s->prof_launches, 0, sizeof(s->prof_launches));
	memset(s->prof_ns, 0, sizeof(s->prof_ns));
	return rc;
}

int pp_sim_prof_export(pp_sim_t *s, const char *path) {
	if (!s || !path) return PP_INVALID_ARGS;
	int rc = prof_collect(s);
	if (rc) return rc;
	FILE *fp = fopen(path, "w");
	if (!fp) return PP_UNABLE_SAVE_STATS;
	double total = 0;
	for (int i = 0; i < KI_COUNT; i++) total += s->prof_ns[i];
	for (int i = 0; i < KI_COUNT; i++)
		fprintf(fp, "\"%s\"\t%llu\t%lg\n", kernel_names[i],
			(unsigned long long)s->prof_ns[i],
			total > 0 ? s->prof_ns[i] / total : 0.0);
	fclose(fp);
	return PP_SUCCESS;
}

int pp_sim_info(const pp_sim_t *s, uint32_t *sort_passes, uint32_t *sort_tile,
	uint32_t *action_tile, uint32_t *grid_move) {
	if (!s) return PP_INVALID_ARGS;
	if (sort_passes) *sort_passes = s->P.passes;
	if (sort_tile) *sort_tile = s->sort_tile;
	if (action_tile) *action_tile = WG;
	if (grid_move) *grid_move = s->grid_move;
	return PP_SUCCESS;
}

int pp_sim_sync(pp_sim_t *s) {
	if (!s) return PP_INVALID_ARGS;
	HIP_OK(hipSetDevice(s->device));
	HIP_OK(hipStreamSynchronize(s->stream));
	return PP_SUCCESS;
}

} /* extern "C" */
