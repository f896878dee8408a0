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
 *                                            O(1) ring update in k_act_scan
 *   move_agent (pp_gpu.cl:744)               k_emit_move (fused with writing
 *                                            the new state), k_move0 after init
 *   cl_ops sbitonic sort (pp_gpu.c:826)      LSD radix sort of (cell, value)
 *                                            pairs: per pass k_sort_up (tile
 *                                            histograms), k_sort_scan (one
 *                                            work-group per digit), k_sort_down
 *                                            (LDS-staged wave64 ranking and
 *                                            scatter); stable, so the order is
 *                                            that of the key cell<<32 | id
 *   find_cell_idx (pp_gpu.cl:834)            segment heads found in k_act
 *   action_agent (pp_gpu.cl:883)             k_act: one lane per occupied cell,
 *                                            shuffled order; k_act_scan: output
 *                                            offsets; k_emit_move: compaction
 *                                            and spawning
 *   reduce_grass1/2, reduce_agent1/2         fused reductions (k_act,
 *   (pp_gpu.cl:485-736)                      k_act_scan) with exact 64-bit sums
 *   stats readback every iteration           device ring of stats rows, one
 *   (pp_gpu.c:766, :840)                     copy per chunk of iterations
 *
 * No kernel waits on another work-group: every cross-group dependency is a
 * kernel boundary, so the step has no spin loops.
 *
 * All arithmetic is integer. The state, its ordering and every random draw are
 * defined so that a single-threaded CPU implementation reproduces the same
 * statistics bit for bit; see the header of each kernel.
 *
 * Device state of one simulation (capacity = A agents, C cells, T sort tiles):
 *   agents (SoA)      cell u32[A], energy i32[A], type u8[A]       9 B/agent
 *   sort pairs        key u32[A], val u32[A], three buffers       24 B/agent
 *   grass             regrow step modulo 2^8 (R <= 191), 2^16 or 2^32,
 *                     1, 2 or 4 B/cell
 *   grass ring        u64[R+1] cells per regrow step (mod R+1)
 *   tile histograms   u32[radix][T], digit totals u32[radix]
 *   action tiles      u32 output count per 1024 sorted positions
 *   control block     counters, live count, step, error
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
constexpr int RADIX_MAX = 2048;     /* digits of at most 11 bits */
constexpr uint32_t DEAD_KEY = 0xFFFFFFFFu;
constexpr uint32_t NO_DIGIT = 0xFFFFFFFFu;
constexpr uint32_t STATS_RING = 4096;
constexpr int M_IPT = 4;            /* agents per thread (move) */
constexpr int A_IPT = 4;            /* sorted positions per thread (action) */
constexpr int A_TILE = WG * A_IPT;
constexpr int A_STAGE = 2 * A_TILE; /* staged outputs per action tile */
/* Grass regrow steps are stored modulo 2^(8 sizeof(GT)); alive cells are
 * re-stamped once every REFRESH iterations (1/REFRESH of the grid per
 * iteration), so every stored value is within R + REFRESH + 1 of the step. */
constexpr uint32_t REFRESH = 64;

/* Packed agent value carried through the sort and the action kernels. */
constexpr uint32_t V_ENERGY = 0x00FFFFFFu; /* energy, 24 bits */
constexpr uint32_t V_WOLF = 1u << 24;      /* type (0 sheep, 1 wolf) */
constexpr uint32_t V_REPRO = 1u << 25;     /* reproduction draw < probability */
constexpr uint32_t V_KILLED = 1u << 26;    /* eaten by a wolf this step */
constexpr uint32_t V_CHILD = 1u << 27;     /* reproduced this step */
constexpr uint32_t V_ODD = 1u << 28;       /* energy before the split was odd */

struct Ctl {
	uint32_t n;        /* agents of the current state */
	uint32_t step;     /* completed iterations */
	int32_t error;     /* 0 or a PP_* code; every kernel is a no-op once set */
	uint32_t err_step; /* iteration at which error was set */
	uint32_t live[2];  /* survivors of the move of iteration s, in [s & 1] */
	uint32_t out_n;    /* agents after actions and births */
	uint32_t pad;
	unsigned long long dead_cells;  /* grass cells with countdown > 0 */
	unsigned long long cd_sum;      /* sum of grass countdowns */
	unsigned long long eaten;       /* grass eaten in this iteration */
	unsigned long long acc[4];      /* sheep, wolves, sheep energy, wolf energy */
	unsigned long long occupied;    /* occupied cells in this iteration */
	unsigned long long agent_steps; /* sum of n over completed iterations */
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
	uint32_t passes;     /* radix passes over the cell index */
	uint32_t digit_bits; /* bits per pass (<= 11) */
	uint32_t sort_tile;  /* positions per sort tile */
	uint32_t hist_cols;  /* tile capacity of the histogram table */
};

/* ---------------------------------------------------------------- helpers */

/* Work-group exclusive scan of one value per thread (NT threads). */
template <int NT>
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
	for (int i = 0; i < NT / WAVE; i++) {
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

/* Work-group sum of K counters into global atomics (one per group). */
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

/* Grass of a cell, stored as its regrow step a (the cell is alive at state
 * s iff a <= s; its countdown is max(0, a - s)) modulo 2^bits of GT. */
template <typename GT>
__device__ __forceinline__ bool grass_alive_at(GT g, uint32_t s, uint32_t R) {
	/* dead iff a in [s + 1, s + R], i.e. (s - a) mod 2^w in [2^w - R, 2^w) */
	const uint32_t w = 8u * sizeof(GT);
	const uint32_t delta = (s - (uint32_t)g) & (w == 32 ? 0xFFFFFFFFu
		: ((1u << w) - 1u));
	const uint32_t dead_from = (w == 32 ? 0u : (1u << w)) - R;
	return delta < dead_from;
}

/* Adds 1 to s_hist[d] for every lane of the wave, merging runs of equal
 * digits in consecutive lanes first (agents arrive ordered by cell, so high
 * digits form long runs). NO_DIGIT lanes count nothing. */
__device__ __forceinline__ void wave_hist_add(uint32_t *s_hist, uint32_t d) {
	const int lane = threadIdx.x & (WAVE - 1);
	const uint32_t prev = __shfl_up(d, 1);
	const bool head = lane == 0 || d != prev;
	const uint64_t heads = __ballot(head);
	if (head && d != NO_DIGIT) {
		const uint64_t after = heads & ~((2ull << lane) - 1ull);
		const uint32_t next = after ? (uint32_t)__builtin_ctzll(after)
			: (uint32_t)WAVE;
		atomicAdd(&s_hist[d], next - (uint32_t)lane);
	}
}

/* One move: Philox block (id, step, MOVE, 0); direction U[0,5) from word 0
 * (0 stay, 1 +x, 2 -x, 3 +y, 4 -y, torus wrap: the table of move_agent,
 * pp_gpu.cl:780-803); reproduction draw U[0,100) from word 1, kept as one bit
 * (draw < probability of the agent's type) for the action phase. Energy -1;
 * the agent dies below 1 (pphpc.nlogo:66-70, pp_gpu.cl:806-815): key
 * DEAD_KEY, dropped by the first sort pass. Returns the key; *val gets the
 * packed value. */
__device__ __forceinline__ uint32_t move_agent(uint32_t id, uint32_t step,
	uint32_t c, int32_t energy, uint32_t wolf, const Params &P, uint32_t *val,
	bool &overflow) {
	const uint4 r = pp_philox(id, step, PP_STREAM_MOVE, 0, P.key0, P.key1);
	const uint32_t dir = pp_bounded(r.x, 5);
	const uint32_t dice = pp_bounded(r.y, 100);
	const int32_t e = energy - 1;
	*val = 0;
	if (e < 1) return DEAD_KEY;
	const uint32_t x = c % P.gx;
	const uint32_t wrap_y = P.C - P.gx;
	uint32_t nc = c;
	if (dir == 1) nc = (x == P.gx - 1) ? c - (P.gx - 1) : c + 1;
	else if (dir == 2) nc = (x == 0) ? c + (P.gx - 1) : c - 1;
	else if (dir == 3) nc = (c >= wrap_y) ? c - wrap_y : c + P.gx;
	else if (dir == 4) nc = (c < P.gx) ? c + wrap_y : c - P.gx;
	const uint32_t prob = wolf ? P.wolves_prob : P.sheep_prob;
	overflow |= e > (int32_t)V_ENERGY;
	*val = ((uint32_t)e & V_ENERGY) | (wolf ? V_WOLF : 0u)
		| (dice < prob ? V_REPRO : 0u);
	return nc;
}

/* ------------------------------------------------------------------- init */

/* Zero the iteration counters, the grass ring and the pass-0 histograms. */
__global__ __launch_bounds__(WG) void k_init_reset(Ctl *ctl,
	unsigned long long *ring, uint32_t ring_n, uint32_t *hist,
	uint32_t hist_words) {
	const uint32_t gtid = blockIdx.x * WG + threadIdx.x;
	for (uint32_t i = gtid; i < ring_n; i += gridDim.x * WG) ring[i] = 0;
	for (uint32_t i = gtid; i < hist_words; i += gridDim.x * WG) hist[i] = 0;
	if (gtid == 0) {
		ctl->n = 0;
		ctl->step = 0;
		ctl->error = 0;
		ctl->err_step = 0xFFFFFFFFu;
		ctl->live[0] = 0;
		ctl->live[1] = 0;
		ctl->out_n = 0;
		ctl->dead_cells = 0;
		ctl->cd_sum = 0;
		ctl->eaten = 0;
		for (int i = 0; i < 4; i++) ctl->acc[i] = 0;
		ctl->occupied = 0;
		ctl->agent_steps = 0;
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
template <typename GT>
__global__ __launch_bounds__(WG) void k_init_cells(GT *grass_at,
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
		grass_at[c] = (GT)g;
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

/* Move of iteration `step` for a state stored as SoA in id order (the state
 * right after init): pairs for the sort plus the pass-0 digit histogram of
 * every sort tile. Each thread moves M_IPT consecutive agents with 16-byte
 * loads and stores. */
__global__ __launch_bounds__(WG) void k_move0(const uint32_t *__restrict__ cell,
	const int32_t *__restrict__ energy, const uint8_t *__restrict__ type,
	uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
	uint32_t *__restrict__ hist, Ctl *ctl, Params P) {
	__shared__ uint32_t s_hist[RADIX_MAX];
	__shared__ unsigned long long s_red[NWAVE][1];
	if (ctl->error) return;
	const uint32_t n = ctl->n, step = ctl->step;
	const int tid = threadIdx.x;
	const uint32_t radix = 1u << P.digit_bits, dmask = radix - 1u;
	unsigned long long live[1] = {0};
	bool overflow = false;
	constexpr uint32_t CHUNK = WG * M_IPT;   /* divides every sort tile */
	for (uint32_t cb = blockIdx.x * CHUNK; cb < n; cb += gridDim.x * CHUNK) {
		for (uint32_t i = tid; i < radix; i += WG) s_hist[i] = 0;
		__syncthreads();
		const uint32_t i0 = cb + tid * M_IPT;
		uint32_t c[M_IPT], t[M_IPT];
		int32_t e[M_IPT];
		if (i0 + M_IPT <= n) {
			const uint4 cv = *reinterpret_cast<const uint4 *>(cell + i0);
			const int4 ev = *reinterpret_cast<const int4 *>(energy + i0);
			const uint32_t tv = *reinterpret_cast<const uint32_t *>(type + i0);
			c[0] = cv.x; c[1] = cv.y; c[2] = cv.z; c[3] = cv.w;
			e[0] = ev.x; e[1] = ev.y; e[2] = ev.z; e[3] = ev.w;
#pragma unroll
			for (int j = 0; j < M_IPT; j++) t[j] = (tv >> (8 * j)) & 0xFFu;
		} else {
#pragma unroll
			for (int j = 0; j < M_IPT; j++) {
				const bool in = i0 + j < n;
				c[j] = in ? cell[i0 + j] : 0;
				e[j] = in ? energy[i0 + j] : 0;
				t[j] = in ? type[i0 + j] : 0;
			}
		}
		uint32_t k[M_IPT], v[M_IPT];
#pragma unroll
		for (int j = 0; j < M_IPT; j++) {
			k[j] = move_agent(i0 + j, step, c[j], e[j], t[j], P, &v[j], overflow);
			if (i0 + j >= n) k[j] = DEAD_KEY;
			live[0] += k[j] != DEAD_KEY;
		}
		if (i0 + M_IPT <= n) {
			*reinterpret_cast<uint4 *>(keys + i0) = make_uint4(k[0], k[1], k[2], k[3]);
			*reinterpret_cast<uint4 *>(vals + i0) = make_uint4(v[0], v[1], v[2], v[3]);
		} else {
#pragma unroll
			for (int j = 0; j < M_IPT; j++)
				if (i0 + j < n) {
					keys[i0 + j] = k[j];
					vals[i0 + j] = v[j];
				}
		}
#pragma unroll
		for (int j = 0; j < M_IPT; j++)
			wave_hist_add(s_hist, k[j] == DEAD_KEY ? NO_DIGIT : (k[j] & dmask));
		__syncthreads();
		const uint32_t tile = cb / P.sort_tile;
		for (uint32_t d = tid; d < radix; d += WG)
			if (s_hist[d]) atomicAdd(&hist[(size_t)d * P.hist_cols + tile], s_hist[d]);
		__syncthreads();
	}
	if (overflow) set_error(ctl, PP_OUT_OF_RESOURCES, step);
	const int lane = tid & (WAVE - 1), w = tid / WAVE;
	const unsigned long long tsum = wave_sum(live[0]);
	if (lane == 0) s_red[w][0] = tsum;
	__syncthreads();
	if (tid == 0) {
		unsigned long long sum = 0;
		for (int i = 0; i < NWAVE; i++) sum += s_red[i][0];
		if (sum) atomicAdd(&ctl->live[step & 1], (uint32_t)sum);
	}
}

/* ------------------------------------------------------------------- sort */

/* Radix sort of the (cell, value) pairs by cell, least significant digit
 * first, digit_bits per pass. Every pass is stable, so survivors end ordered
 * by (cell, agent id) -- the key (cell << 32 | agent id) -- and the first pass
 * drops dead agents (DEAD_KEY). A pass is three kernels:
 *   k_sort_up    digit histogram of every tile (pass 0: built by the move)
 *   k_sort_scan  per digit, exclusive prefix over tiles and the digit total
 *   k_sort_down  per tile: rank in LDS, scatter to bucket start + tile prefix
 * Tile t covers positions [t * sort_tile, (t + 1) * sort_tile). The table
 * hist[d * hist_cols + t] holds tile counts, then tile offsets. */

__device__ __forceinline__ uint32_t pass_len(const Ctl *ctl, uint32_t pass,
	uint32_t step) {
	return pass == 0 ? ctl->n : ctl->live[step & 1];
}

template <int KPT>
__global__ __launch_bounds__(WG) void k_sort_up(const uint32_t *__restrict__ kin,
	uint32_t *__restrict__ hist, const Ctl *ctl, uint32_t pass, Params P) {
	constexpr uint32_t TILE = WG * KPT;
	__shared__ uint32_t s_hist[RADIX_MAX];
	if (ctl->error) return;
	const uint32_t n = pass_len(ctl, pass, ctl->step);
	const uint32_t ntiles = (n + TILE - 1) / TILE;
	const uint32_t shift = pass * P.digit_bits;
	const uint32_t radix = 1u << P.digit_bits, dmask = radix - 1u;
	const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
	for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		for (uint32_t i = tid; i < radix; i += WG) s_hist[i] = 0;
		__syncthreads();
		const uint32_t base = tile * TILE + w * (WAVE * KPT);
#pragma unroll
		for (int j = 0; j < KPT; j++) {
			const uint32_t idx = base + j * WAVE + lane;
			const uint32_t k = idx < n ? kin[idx] : DEAD_KEY;
			wave_hist_add(s_hist, k == DEAD_KEY ? NO_DIGIT : ((k >> shift) & dmask));
		}
		__syncthreads();
		for (uint32_t d = tid; d < radix; d += WG)
			hist[(size_t)d * P.hist_cols + tile] = s_hist[d];
		__syncthreads();
	}
}

/* One work-group per digit: hist[d][0..ntiles) -> exclusive prefix, and the
 * digit's total. Each thread sums a contiguous run of tiles. The pass-0 scan
 * also clears the action tiles' counters for this iteration and re-stamps
 * one grass slice. */
template <typename GT>
__global__ __launch_bounds__(WG) void k_sort_scan(uint32_t *__restrict__ hist,
	uint32_t *__restrict__ totals, const Ctl *ctl, uint32_t pass, Params P,
	uint32_t *__restrict__ act_cnt, uint32_t act_words, GT *grass_at) {
	__shared__ uint32_t s_tmp[NWAVE];
	if (ctl->error) return;
	if (pass == 0) {
		/* The action tiles' counters of this iteration start at zero. */
		for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < act_words;
			i += gridDim.x * WG)
			act_cnt[i] = 0;
		/* Re-stamp the alive cells of slice (s mod REFRESH) with s, which
		 * keeps their stored regrow step within the modular window; the
		 * state does not change (alive stays alive, countdown 0). */
		const uint32_t s = ctl->step;
		const uint32_t per = (P.C + REFRESH - 1) / REFRESH;
		const uint32_t c0 = (s % REFRESH) * per;
		const uint32_t c1 = min(c0 + per, P.C);
		for (uint32_t c = c0 + blockIdx.x * WG + threadIdx.x; c < c1;
			c += gridDim.x * WG) {
			const GT g = grass_at[c];
			if (grass_alive_at<GT>(g, s, P.R) && g != (GT)s) grass_at[c] = (GT)s;
		}
	}
	const uint32_t n = pass_len(ctl, pass, ctl->step);
	const uint32_t ntiles = (n + P.sort_tile - 1) / P.sort_tile;
	const uint32_t radix = 1u << P.digit_bits;
	const uint32_t per = (ntiles + WG - 1) / WG;
	const uint32_t t0 = threadIdx.x * per;
	const uint32_t t1 = min(t0 + per, ntiles);
	for (uint32_t d = blockIdx.x; d < radix; d += gridDim.x) {
		uint32_t *col = hist + (size_t)d * P.hist_cols;
		uint32_t sum = 0;
		for (uint32_t t = t0; t < t1; t++) sum += col[t];
		uint32_t total;
		uint32_t run = wg_excl_scan<WG>(sum, s_tmp, total);
		for (uint32_t t = t0; t < t1; t++) {
			const uint32_t v = col[t];
			col[t] = run;
			run += v;
		}
		if (threadIdx.x == 0) totals[d] = total;
	}
}

/* Per tile: load KPT keys per thread (wave w owns [tile*TILE + w*64*KPT,
 * +64*KPT), element j of lane l at offset j*64 + l), rank them inside each
 * wave with ballots over the digit bits (order (j, lane) = input order, so the
 * ranking is stable), combine the waves, stage the tile in digit order in
 * LDS and write it to bucket start + tile prefix. */
template <int KPT, int RBITS>
__global__ __launch_bounds__(WG) void k_sort_down(
	const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
	uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
	const uint32_t *__restrict__ hist, const uint32_t *__restrict__ totals,
	const Ctl *ctl, uint32_t pass, Params P) {
	constexpr uint32_t TILE = WG * KPT;
	constexpr uint32_t R = 1u << RBITS;
	constexpr uint32_t DPT = R / WG < 1 ? 1 : R / WG;  /* digits per thread */
	__shared__ uint32_t s_keys[TILE];
	__shared__ uint32_t s_vals[TILE];
	__shared__ uint32_t s_wh[NWAVE][R + 1];
	__shared__ uint32_t s_bstart[R];
	__shared__ int32_t s_dst[R];
	__shared__ uint32_t s_tstart[R];
	__shared__ uint32_t s_tmp[NWAVE];

	if (ctl->error) return;
	const uint32_t n = pass_len(ctl, pass, ctl->step);
	const uint32_t ntiles = (n + TILE - 1) / TILE;
	const uint32_t shift = pass * P.digit_bits;
	const uint32_t radix = 1u << P.digit_bits, dmask = radix - 1u;
	const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
	const uint64_t lt_mask = (1ull << lane) - 1ull;

	/* Bucket starts: exclusive scan over digits of the digit totals. Thread
	 * tid owns digits [tid*DPT, tid*DPT + DPT). */
	{
		uint32_t loc[DPT], sum = 0;
#pragma unroll
		for (uint32_t i = 0; i < DPT; i++) {
			const uint32_t d = tid * DPT + i;
			loc[i] = d < radix ? totals[d] : 0u;
			sum += loc[i];
		}
		uint32_t tot;
		uint32_t run = wg_excl_scan<WG>(sum, s_tmp, tot);
#pragma unroll
		for (uint32_t i = 0; i < DPT; i++) {
			const uint32_t d = tid * DPT + i;
			if (d < R) s_bstart[d] = run;
			run += loc[i];
		}
	}

	for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		for (uint32_t i = tid; i < NWAVE * (R + 1); i += WG) (&s_wh[0][0])[i] = 0;
		__syncthreads();
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
			dig[j] = (k == DEAD_KEY) ? R : ((k >> shift) & dmask);
		}
#pragma unroll
		for (int j = 0; j < KPT; j++) {
			const uint32_t d = dig[j];
			uint64_t peers = ~0ull;
#pragma unroll
			for (int b = 0; b <= RBITS; b++) {
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

		/* Per digit: exclusive offsets of the waves, tile count; then the
		 * tile-local start of each digit (exclusive scan over digits). */
		uint32_t cnt_d[DPT], sum = 0;
#pragma unroll
		for (uint32_t i = 0; i < DPT; i++) {
			const uint32_t d = tid * DPT + i;
			uint32_t acc = 0;
#pragma unroll
			for (int ww = 0; ww < NWAVE; ww++) {
				const uint32_t c = s_wh[ww][d];
				s_wh[ww][d] = acc;
				acc += c;
			}
			cnt_d[i] = acc;
			sum += acc;
		}
		uint32_t tile_live;
		uint32_t run = wg_excl_scan<WG>(sum, s_tmp, tile_live);
#pragma unroll
		for (uint32_t i = 0; i < DPT; i++) {
			const uint32_t d = tid * DPT + i;
			s_tstart[d] = run;
			const uint32_t off = d < radix ? hist[(size_t)d * P.hist_cols + tile] : 0u;
			s_dst[d] = (int32_t)(s_bstart[d] + off) - (int32_t)run;
			run += cnt_d[i];
		}
		__syncthreads();
#pragma unroll
		for (int j = 0; j < KPT; j++) {
			const uint32_t d = dig[j];
			if (d < R) {
				const uint32_t q = s_tstart[d] + s_wh[w][d] + loc[j];
				s_keys[q] = key[j];
				s_vals[q] = val[j];
			}
		}
		__syncthreads();
		for (uint32_t q = tid; q < tile_live; q += WG) {
			const uint32_t k = s_keys[q];
			const uint32_t d = (k >> shift) & dmask;
			const uint32_t dst = (uint32_t)(s_dst[d] + (int32_t)q);
			kout[dst] = k;
			vout[dst] = s_vals[q];
		}
		__syncthreads();
	}
}

/* ----------------------------------------------------------------- action */

/* Values of one cell's segment: positions of the current tile live in LDS,
 * the few positions of a segment running into the next tile in global
 * memory (only the lane that owns the segment touches them). */
struct SegVals {
	uint32_t *s_val;
	uint32_t *g_val;
	__device__ __forceinline__ uint32_t get(uint32_t q) const {
		return q < (uint32_t)A_TILE ? s_val[q] : g_val[q];
	}
	__device__ __forceinline__ void set(uint32_t q, uint32_t v) const {
		if (q < (uint32_t)A_TILE) s_val[q] = v;
		else g_val[q] = v;
	}
};

/* End (exclusive, tile-local) of the segment of cell c that starts at lp. */
__device__ __forceinline__ uint32_t seg_end(const uint32_t *s_key,
	const uint32_t *__restrict__ keys, uint32_t lp, uint32_t cnt,
	uint32_t base, uint32_t L, uint32_t c) {
	uint32_t e = lp + 1;
	while (e < cnt && s_key[1 + e] == c) e++;
	if (e == cnt) {
		uint32_t g = base + e;
		while (g < L && keys[g] == c) g++;
		e = g - base;
	}
	return e;
}

/* Loads positions [base, base + A_TILE) of the sorted keys (s_key[1..], with
 * s_key[0] the key just before the tile) and optionally the values. */
__device__ __forceinline__ uint32_t load_tile(const uint32_t *__restrict__ keys,
	const uint32_t *vals, uint32_t *s_key, uint32_t *s_val, uint32_t base,
	uint32_t L) {
	const uint32_t cnt = min((uint32_t)A_TILE, L - base);
	for (uint32_t i = threadIdx.x; i < (uint32_t)A_TILE; i += WG) {
		const bool in = i < cnt;
		s_key[1 + i] = in ? keys[base + i] : DEAD_KEY;
		if (s_val) s_val[i] = in ? vals[base + i] : 0u;
	}
	if (threadIdx.x == 0) s_key[0] = base ? keys[base - 1] : DEAD_KEY;
	__syncthreads();
	return cnt;
}

/* Agent actions for the agents sorted by cell (ascending agent id within a
 * cell). A work-group stages A_TILE sorted positions in LDS; the thread owning
 * a position that starts a run of equal cells (the segment that find_cell_idx,
 * pp_gpu.cl:834-867, delimits) acts for the whole cell in sequence:
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
 * The permuted, updated values are written back in place, so the next state
 * is the sorted positions in order, each emitting its survivor then its
 * newborn; act_cnt[t] (zeroed by the pass-0 scan) receives the number of
 * agents the positions of tile t emit. Counts and energy sums of the new state
 * are reduced here. This kernel also clears the pass-0 histogram table that
 * k_emit_move fills for the next iteration. */
template <typename GT>
__global__ __launch_bounds__(WG) void k_act(const uint32_t *__restrict__ keys,
	uint32_t *vals, GT *grass_at, uint32_t *act_cnt,
	uint32_t *__restrict__ hist, uint32_t hist_words, Ctl *ctl, Params P) {
	__shared__ uint32_t s_key[A_TILE + 1];
	__shared__ uint32_t s_val[A_TILE];
	__shared__ uint32_t s_slow_lp[A_TILE];
	__shared__ GT s_slow_g[A_TILE];
	__shared__ uint32_t s_nslow;
	__shared__ uint32_t s_tmp[NWAVE];
	__shared__ unsigned long long s_red[NWAVE][6];

	if (ctl->error) return;
	const int tid = threadIdx.x;
	for (uint32_t i = blockIdx.x * WG + tid; i < hist_words; i += gridDim.x * WG)
		hist[i] = 0;
	const uint32_t step = ctl->step;
	const uint32_t L = ctl->live[step & 1];
	const uint32_t ntiles = (L + A_TILE - 1) / A_TILE;
	const uint32_t alive_at = step + 1;
	const GT eaten_until = (GT)(step + 1 + P.R);
	unsigned long long acc[6] = {0, 0, 0, 0, 0, 0};
	bool overflow = false;

	for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const uint32_t base = tile * A_TILE;
		const uint32_t cnt = load_tile(keys, vals, s_key, s_val, base, L);
		const SegVals sv{s_val, vals + base};
		/* Heads of this thread's positions; their grass loads are issued
		 * together before any segment is processed. */
		uint32_t cj[A_IPT];
		GT gj[A_IPT];
		bool hj[A_IPT];
#pragma unroll
		for (int j = 0; j < A_IPT; j++) {
			const uint32_t lp = tid * A_IPT + j;
			cj[j] = s_key[1 + lp];
			hj[j] = lp < cnt && s_key[lp] != cj[j];
			gj[j] = hj[j] ? grass_at[cj[j]] : (GT)0;
		}
		/* Cells with one agent (most of them) are handled in place; heads of
		 * longer segments are listed and then processed by consecutive
		 * lanes, so the long serial path runs once per wave, not once per
		 * position of every lane. Segments are independent: the order of the
		 * list does not change any result. */
		if (tid == 0) s_nslow = 0;
		__syncthreads();
		uint32_t nout = 0;
#pragma unroll
		for (int j = 0; j < A_IPT; j++) {
			if (!hj[j]) continue;
			const uint32_t lp = tid * A_IPT + j;
			const uint32_t c = cj[j];
			if (lp + 1 < cnt && s_key[2 + lp] != c) {
				/* A cell with one agent (most of them): nothing to shuffle, a
				 * lone wolf finds no sheep, a lone agent is never eaten. */
				const uint32_t v = s_val[lp];
				int32_t e = (int32_t)(v & V_ENERGY);
				const uint32_t t = (v & V_WOLF) ? 1u : 0u;
				uint32_t thr = P.wolves_thr;
				if (!t) {
					thr = P.sheep_thr;
					if (grass_alive_at<GT>(gj[j], alive_at, P.R)) {
						e += (int32_t)P.sheep_gain;
						grass_at[c] = eaten_until;
						acc[4] += 1;
					}
				}
				uint32_t nv = v & V_WOLF;
				uint32_t ce = 0;
				if (e > (int32_t)thr && (v & V_REPRO)) {
					nv |= V_CHILD | ((e & 1) ? V_ODD : 0u);
					ce = (uint32_t)(e / 2);
					e -= e / 2;
				}
				overflow |= e > (int32_t)V_ENERGY;
				s_val[lp] = nv | ((uint32_t)e & V_ENERGY);
				const uint32_t kids = (nv & V_CHILD) ? 1u : 0u;
				acc[5] += 1;
				acc[t] += 1 + kids;
				acc[2 + t] += (uint32_t)e + ce;
				nout += 1 + kids;
				continue;
			}
			const uint32_t slot = atomicAdd(&s_nslow, 1u);
			s_slow_lp[slot] = lp;
			s_slow_g[slot] = gj[j];
		}
		__syncthreads();
		const uint32_t nslow = s_nslow;
#pragma unroll 1
		for (uint32_t si = tid; si < nslow; si += WG) {
			const uint32_t lp = s_slow_lp[si];
			const uint32_t c = s_key[1 + lp];
			const GT gslow = s_slow_g[si];
			const uint32_t end = seg_end(s_key, keys, lp, cnt, base, L, c);
			const uint32_t k = end - lp;
			acc[5] += 1;
			if (k > 1) {
				uint4 r = make_uint4(0, 0, 0, 0);
				for (uint32_t i = k - 1, m = 0; i > 0; i--, m++) {
					if ((m & 3u) == 0)
						r = pp_philox(c, step, PP_STREAM_SHUFFLE, m >> 2, P.key0,
							P.key1);
					const uint32_t word = (m & 3u) == 0 ? r.x
						: (m & 3u) == 1 ? r.y : (m & 3u) == 2 ? r.z : r.w;
					const uint32_t jj = pp_bounded(word, i + 1);
					if (jj != i) {
						const uint32_t a = sv.get(lp + i);
						sv.set(lp + i, sv.get(lp + jj));
						sv.set(lp + jj, a);
					}
				}
			}
			bool grass = grass_alive_at<GT>(gslow, alive_at, P.R);
			bool ate = false;
			for (uint32_t q = lp; q < end; q++) {
				const uint32_t v = sv.get(q);
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
					for (uint32_t o = lp; o < end; o++) {
						const uint32_t u = sv.get(o);
						if (!(u & (V_WOLF | V_KILLED))) {
							sv.set(o, u | V_KILLED);
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
				sv.set(q, nv | ((uint32_t)e & V_ENERGY));
			}
			if (ate) {
				grass_at[c] = eaten_until;
				acc[4] += 1;
			}
			for (uint32_t q = lp; q < end; q++) {
				const uint32_t v = sv.get(q);
				const uint32_t e = v & V_ENERGY;
				const uint32_t t = (v & V_WOLF) ? 1u : 0u;
				const uint32_t outs = ((v & V_KILLED) == 0) + ((v & V_CHILD) != 0);
				if (!(v & V_KILLED)) {
					acc[t] += 1;
					acc[2 + t] += e;
				}
				if (v & V_CHILD) {
					acc[t] += 1;
					acc[2 + t] += (v & V_ODD) ? e - 1 : e;
				}
				if (q < (uint32_t)A_TILE) nout += outs;
				else if (outs) atomicAdd(&act_cnt[tile + q / A_TILE], outs);
			}
		}
		__syncthreads();
		/* Write back the positions this tile owns (not those of the previous
		 * tile's last segment, which its lane wrote). */
		for (uint32_t i = tid; i < cnt; i += WG)
			if (s_key[1 + i] != s_key[0]) vals[base + i] = s_val[i];
		uint32_t total;
		(void)wg_excl_scan<WG>(nout, s_tmp, total);
		if (tid == 0 && total) atomicAdd(&act_cnt[tile], total);
	}
	if (overflow) set_error(ctl, PP_OUT_OF_RESOURCES, step);
	unsigned long long *dst[6] = {&ctl->acc[0], &ctl->acc[1], &ctl->acc[2],
		&ctl->acc[3], &ctl->eaten, &ctl->occupied};
	wg_add<6>(acc, s_red, dst);
}

/* One work-group: exclusive scan of the action tiles' output counts (the
 * output offsets), then the end of iteration s -> s+1. Grass statistics come
 * from the regrow-step encoding without touching the grid: with D dead cells
 * and countdown sum S at state s, every dead countdown drops by one (S -= D),
 * the cells whose regrow step is s+1 come alive (ring slot (s+1) mod (R+1)),
 * and the E cells eaten this iteration restart at R (slot s mod (R+1)):
 * exactly the grass kernel (pp_gpu.cl:448-476) plus the sums of
 * reduce_grass1/2 (pp_gpu.cl:485-583). Then the statistics row s+1 is
 * written, the capacity checked (pp_gpu.c:847-852) and the per-iteration
 * counters cleared. */
constexpr int SCAN_WG = 1024;
constexpr int SCAN_UNROLL = 8;

__global__ __launch_bounds__(SCAN_WG) void k_act_scan(
	uint32_t *__restrict__ act_cnt, Ctl *ctl, unsigned long long *ring,
	pp_stats64_t *stats, Params P) {
	__shared__ uint32_t s_tmp[SCAN_WG / WAVE];
	if (ctl->error) return;
	const uint32_t step = ctl->step;
	const uint32_t L = ctl->live[step & 1];
	const uint32_t nt = (L + A_TILE - 1) / A_TILE;
	const uint32_t per = (nt + SCAN_WG - 1) / SCAN_WG;
	const uint32_t t0 = threadIdx.x * per, t1 = min(t0 + per, nt);
	uint32_t sum = 0;
	for (uint32_t t = t0; t < t1; t += SCAN_UNROLL) {
		uint32_t v[SCAN_UNROLL];
#pragma unroll
		for (int u = 0; u < SCAN_UNROLL; u++)
			v[u] = t + u < t1 ? act_cnt[t + u] : 0u;
#pragma unroll
		for (int u = 0; u < SCAN_UNROLL; u++) sum += v[u];
	}
	uint32_t total;
	uint32_t run = wg_excl_scan<SCAN_WG>(sum, s_tmp, total);
	for (uint32_t t = t0; t < t1; t += SCAN_UNROLL) {
		uint32_t v[SCAN_UNROLL];
#pragma unroll
		for (int u = 0; u < SCAN_UNROLL; u++)
			v[u] = t + u < t1 ? act_cnt[t + u] : 0u;
#pragma unroll
		for (int u = 0; u < SCAN_UNROLL; u++) {
			if (t + u < t1) act_cnt[t + u] = run;
			run += v[u];
		}
	}
	if (threadIdx.x != 0) return;
	const uint32_t s = step;
	if (total > P.cap) {
		set_error(ctl, PP_OUT_OF_RESOURCES, s);
		return;
	}
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
	ctl->sum_live += L;
	ctl->sum_out += total;
	ctl->sum_occ += ctl->occupied;
	ctl->sum_eaten += E;
	ctl->n = total;
	ctl->out_n = total;
	for (int i = 0; i < 4; i++) ctl->acc[i] = 0;
	ctl->eaten = 0;
	ctl->occupied = 0;
	ctl->live[(s + 1) & 1] = 0;
	ctl->step = s + 1;
}

/* Emits the new state and moves it for the next iteration. Sorted position
 * p emits its agent unless eaten, then its newborn if any; the agents' ids in
 * the new state are these outputs in position order (tile offset from
 * k_act_scan plus a work-group scan). Each output is staged in LDS, then
 * written with coalesced stores to the SoA state and immediately moved for
 * iteration step (= s+1): its (key, value) pair goes to slot id and its
 * pass-0 digit is counted for the sort tile holding id. The outputs of one
 * action tile (at most 2 x A_TILE) span at most two sort tiles. */
template <int RBITS>
__global__ __launch_bounds__(WG) void k_emit_move(
	const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
	const uint32_t *__restrict__ act_off, uint32_t *__restrict__ ocell,
	int32_t *__restrict__ oenergy, uint8_t *__restrict__ otype,
	uint32_t *__restrict__ mkeys, uint32_t *__restrict__ mvals,
	uint32_t *__restrict__ hist, Ctl *ctl, Params P) {
	constexpr uint32_t R = 1u << RBITS;
	__shared__ uint32_t s_oc[A_STAGE];
	__shared__ uint32_t s_ov[A_STAGE];
	__shared__ uint32_t s_h[2 * R];
	__shared__ uint32_t s_tmp[NWAVE];
	__shared__ unsigned long long s_red[NWAVE][1];

	if (ctl->error) return;
	const int tid = threadIdx.x;
	const uint32_t step = ctl->step;            /* iteration being moved */
	const uint32_t L = ctl->live[(step - 1) & 1];
	const uint32_t ntiles = (L + A_TILE - 1) / A_TILE;
	const uint32_t dmask = (1u << P.digit_bits) - 1u;
	unsigned long long live[1] = {0};
	bool overflow = false;

	for (uint32_t i = tid; i < 2 * R; i += WG) s_h[i] = 0;
	for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
		const uint32_t base = tile * A_TILE;
		const uint32_t goff = act_off[tile];
		const uint32_t p0 = base + tid * A_IPT;
		uint32_t c[A_IPT], v[A_IPT];
		if (p0 + A_IPT <= L) {
			const uint4 kv = *reinterpret_cast<const uint4 *>(keys + p0);
			const uint4 vv = *reinterpret_cast<const uint4 *>(vals + p0);
			c[0] = kv.x; c[1] = kv.y; c[2] = kv.z; c[3] = kv.w;
			v[0] = vv.x; v[1] = vv.y; v[2] = vv.z; v[3] = vv.w;
		} else {
#pragma unroll
			for (int j = 0; j < A_IPT; j++) {
				const bool in = p0 + j < L;
				c[j] = in ? keys[p0 + j] : 0u;
				v[j] = in ? vals[p0 + j] : V_KILLED;
			}
		}
		uint32_t nout = 0;
#pragma unroll
		for (int j = 0; j < A_IPT; j++)
			nout += ((v[j] & V_KILLED) == 0) + ((v[j] & V_CHILD) != 0);
		uint32_t total;
		uint32_t o = wg_excl_scan<WG>(nout, s_tmp, total);
#pragma unroll
		for (int j = 0; j < A_IPT; j++) {
			const uint32_t e = v[j] & V_ENERGY;
			const uint32_t t = (v[j] & V_WOLF) ? 1u : 0u;
			if (!(v[j] & V_KILLED)) {
				s_oc[o] = c[j];
				s_ov[o] = e | (t << 31);
				o++;
			}
			if (v[j] & V_CHILD) {
				s_oc[o] = c[j];
				s_ov[o] = ((v[j] & V_ODD) ? e - 1 : e) | (t << 31);
				o++;
			}
		}
		__syncthreads();
		const uint32_t to0 = goff / P.sort_tile;
		for (uint32_t i0 = 0; i0 < total; i0 += WG) {
			const uint32_t i = i0 + tid;
			uint32_t d = NO_DIGIT;
			if (i < total) {
				const uint32_t id = goff + i;
				const uint32_t cc = s_oc[i], ov = s_ov[i];
				const uint32_t en = ov & V_ENERGY, t = ov >> 31;
				ocell[id] = cc;
				oenergy[id] = (int32_t)en;
				otype[id] = (uint8_t)t;
				uint32_t mv;
				const uint32_t mk = move_agent(id, step, cc, (int32_t)en, t, P,
					&mv, overflow);
				mkeys[id] = mk;
				mvals[id] = mv;
				if (mk != DEAD_KEY) {
					live[0] += 1;
					d = (id / P.sort_tile - to0) * R + (mk & dmask);
				}
			}
			wave_hist_add(s_h, d);
		}
		__syncthreads();
		for (uint32_t i = tid; i < 2 * R; i += WG) {
			const uint32_t h = s_h[i];
			if (h) {
				atomicAdd(&hist[(size_t)(i % R) * P.hist_cols + to0 + i / R], h);
				s_h[i] = 0;
			}
		}
		__syncthreads();
	}
	if (overflow) set_error(ctl, PP_OUT_OF_RESOURCES, step);
	const int lane = tid & (WAVE - 1), w = tid / WAVE;
	const unsigned long long tsum = wave_sum(live[0]);
	if (lane == 0) s_red[w][0] = tsum;
	__syncthreads();
	if (tid == 0) {
		unsigned long long sum = 0;
		for (int i = 0; i < NWAVE; i++) sum += s_red[i][0];
		if (sum) atomicAdd(&ctl->live[step & 1], (uint32_t)sum);
	}
}

/* ------------------------------------------------------------ host layer */

enum KernelId {
	KI_INIT_RESET,
	KI_INIT_CELLS,
	KI_INIT_AGENTS,
	KI_INIT_DONE,
	KI_MOVE0,
	KI_SORT_UP,
	KI_SORT_SCAN,
	KI_SORT_DOWN,
	KI_ACT,
	KI_ACT_SCAN,
	KI_EMIT_MOVE,
	KI_COUNT
};

const char *const kernel_names[KI_COUNT] = {
	"init_reset", "init_cells", "init_agents", "init_done", "move0",
	"sort_up", "sort_scan", "sort_down", "act", "act_scan", "emit_move",
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
	uint32_t *keys[3];
	uint32_t *vals[3];
	void *grass_at;
	uint32_t grass_bytes;     /* 1, 2 or 4 bytes per cell */
	unsigned long long *ring;
	uint32_t ring_n;
	uint32_t *hist;
	uint32_t hist_words;
	uint32_t *totals;
	uint32_t *act_cnt;
	uint32_t act_words;
	Ctl *ctl;
	pp_stats64_t *stats_dev;
	pp_stats64_t *stats_host; /* pinned, STATS_RING rows */
	Ctl *ctl_host;            /* pinned mirror */

	int kpt;
	int rbits;                /* template radix bits of the sort kernels */
	uint32_t grid_move, grid_up, grid_down, grid_act, grid_emit, grid_cells;
	uint32_t grid_agents;

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

/* Pair buffers: the move writes buffer 0; pass p reads buffer 0 (p = 0) or
 * 1 + ((p - 1) & 1) and writes 1 + (p & 1). */
static int pass_src(uint32_t p) { return p == 0 ? 0 : 1 + (int)((p - 1) & 1); }
static int pass_dst(uint32_t p) { return 1 + (int)(p & 1); }

template <int KPT, int RBITS>
static void launch_sort(pp_sim_t *s, uint32_t pass) {
	PendingEvent pe;
	const int src = pass_src(pass), dst = pass_dst(pass);
	if (pass > 0) {
		prof_begin(s, KI_SORT_UP, pe);
		k_sort_up<KPT><<<s->grid_up, WG, 0, s->stream>>>(s->keys[src], s->hist,
			s->ctl, pass, s->P);
		prof_end(s, pe);
	}
	prof_begin(s, KI_SORT_SCAN, pe);
	const uint32_t nscan = 1u << s->P.digit_bits;
	if (s->grass_bytes == 1)
		k_sort_scan<uint8_t><<<nscan, WG, 0, s->stream>>>(s->hist, s->totals,
			s->ctl, pass, s->P, s->act_cnt, s->act_words,
			(uint8_t *)s->grass_at);
	else if (s->grass_bytes == 2)
		k_sort_scan<uint16_t><<<nscan, WG, 0, s->stream>>>(s->hist, s->totals,
			s->ctl, pass, s->P, s->act_cnt, s->act_words,
			(uint16_t *)s->grass_at);
	else
		k_sort_scan<uint32_t><<<nscan, WG, 0, s->stream>>>(s->hist, s->totals,
			s->ctl, pass, s->P, s->act_cnt, s->act_words,
			(uint32_t *)s->grass_at);
	prof_end(s, pe);
	prof_begin(s, KI_SORT_DOWN, pe);
	k_sort_down<KPT, RBITS><<<s->grid_down, WG, 0, s->stream>>>(s->keys[src],
		s->vals[src], s->keys[dst], s->vals[dst], s->hist, s->totals, s->ctl,
		pass, s->P);
	prof_end(s, pe);
}

static void launch_sort_pass(pp_sim_t *s, uint32_t pass) {
	if (s->kpt == 8) {
		if (s->rbits == 11) launch_sort<8, 11>(s, pass);
		else launch_sort<8, 8>(s, pass);
	} else {
		if (s->rbits == 11) launch_sort<16, 11>(s, pass);
		else launch_sort<16, 8>(s, pass);
	}
}

/* Enqueue one iteration (no synchronisation): the sort of the pairs the
 * previous iteration's emit (or the move after init) produced, the actions,
 * the scan + end of iteration, and the emit + move of the next iteration. */
static int enqueue_step(pp_sim_t *s) {
	PendingEvent pe;
	for (uint32_t pass = 0; pass < s->P.passes; pass++) launch_sort_pass(s, pass);
	const int sorted = pass_dst(s->P.passes - 1);
	prof_begin(s, KI_ACT, pe);
	if (s->grass_bytes == 1)
		k_act<uint8_t><<<s->grid_act, WG, 0, s->stream>>>(s->keys[sorted],
			s->vals[sorted], (uint8_t *)s->grass_at, s->act_cnt, s->hist,
			s->hist_words, s->ctl, s->P);
	else if (s->grass_bytes == 2)
		k_act<uint16_t><<<s->grid_act, WG, 0, s->stream>>>(s->keys[sorted],
			s->vals[sorted], (uint16_t *)s->grass_at, s->act_cnt, s->hist,
			s->hist_words, s->ctl, s->P);
	else
		k_act<uint32_t><<<s->grid_act, WG, 0, s->stream>>>(s->keys[sorted],
			s->vals[sorted], (uint32_t *)s->grass_at, s->act_cnt, s->hist,
			s->hist_words, s->ctl, s->P);
	prof_end(s, pe);
	prof_begin(s, KI_ACT_SCAN, pe);
	k_act_scan<<<1, SCAN_WG, 0, s->stream>>>(s->act_cnt, s->ctl, s->ring,
		s->stats_dev, s->P);
	prof_end(s, pe);
	prof_begin(s, KI_EMIT_MOVE, pe);
	if (s->rbits == 11)
		k_emit_move<11><<<s->grid_emit, WG, 0, s->stream>>>(s->keys[sorted],
			s->vals[sorted], s->act_cnt, s->cell, s->energy, s->type,
			s->keys[0], s->vals[0], s->hist, s->ctl, s->P);
	else
		k_emit_move<8><<<s->grid_emit, WG, 0, s->stream>>>(s->keys[sorted],
			s->vals[sorted], s->act_cnt, s->cell, s->energy, s->type,
			s->keys[0], s->vals[0], s->hist, s->ctl, s->P);
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
	for (int i = 0; i < 3; i++) {
		(void)hipFree(s->keys[i]);
		(void)hipFree(s->vals[i]);
	}
	(void)hipFree(s->grass_at);
	(void)hipFree(s->ring);
	(void)hipFree(s->hist);
	(void)hipFree(s->totals);
	(void)hipFree(s->act_cnt);
	(void)hipFree(s->ctl);
	(void)hipFree(s->stats_dev);
	if (s->stats_host) (void)hipHostFree(s->stats_host);
	if (s->ctl_host) (void)hipHostFree(s->ctl_host);
	if (s->stream) (void)hipStreamDestroy(s->stream);
	delete s;
}

static int sim_alloc(pp_sim_t *s) {
	const uint64_t A = s->cap, C = s->P.C;
	/* Vector accesses of the move read/write 16 bytes at multiples of 4. */
	const uint64_t A4 = (A + 3) & ~3ull;
	HIP_OK(hipMalloc(&s->cell, A4 * 4));
	HIP_OK(hipMalloc(&s->energy, A4 * 4));
	HIP_OK(hipMalloc(&s->type, A4));
	for (int i = 0; i < 3; i++) {
		HIP_OK(hipMalloc(&s->keys[i], A4 * 4));
		HIP_OK(hipMalloc(&s->vals[i], A4 * 4));
	}
	HIP_OK(hipMalloc(&s->grass_at, C * s->grass_bytes));
	s->ring_n = s->P.R + 1;
	HIP_OK(hipMalloc(&s->ring, (size_t)s->ring_n * 8));
	s->hist_words = (1u << s->P.digit_bits) * s->P.hist_cols;
	HIP_OK(hipMalloc(&s->hist, (size_t)s->hist_words * 4));
	HIP_OK(hipMalloc(&s->totals, RADIX_MAX * 4));
	s->act_words = ceil_div(A, A_TILE) + 1;
	HIP_OK(hipMalloc(&s->act_cnt, (size_t)s->act_words * 4));
	HIP_OK(hipMalloc(&s->ctl, sizeof(Ctl)));
	HIP_OK(hipMalloc(&s->stats_dev, STATS_RING * sizeof(pp_stats64_t)));
	HIP_OK(hipHostMalloc(&s->stats_host, STATS_RING * sizeof(pp_stats64_t),
		hipHostMallocDefault));
	HIP_OK(hipHostMalloc(&s->ctl_host, sizeof(Ctl), hipHostMallocDefault));
	HIP_OK(hipMemset(s->ctl, 0, sizeof(Ctl)));
	HIP_OK(hipDeviceSynchronize());
	memset(s->ctl_host, 0, sizeof(Ctl));
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
	/* Output offsets and ids are 32-bit, and an iteration can at most double
	 * the survivors. */
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
	/* Radix passes over the significant bits of the largest cell index, the
	 * bits spread evenly: up to 22 bits in passes of at most 11 bits (2048
	 * buckets; 1600 x 1600 = 2 x 11), wider indices in passes of at most 8
	 * bits (12800 x 12800 = 4 x 7), where longer runs per bucket keep the
	 * scatter coalesced. */
	uint32_t bits = 0;
	while (bits < 32 && ((C - 1) >> bits) != 0) bits++;
	if (bits == 0) bits = 1;
	const uint32_t max_bits = bits <= 22 ? 11 : 8;
	P.passes = (bits + max_bits - 1) / max_bits;
	P.digit_bits = (bits + P.passes - 1) / P.passes;
	s->rbits = P.digit_bits > 8 ? 11 : 8;

	/* Narrowest grass storage whose modular window holds R + REFRESH + 1. */
	s->grass_bytes = P.R + REFRESH + 1 <= 256 ? 1
		: P.R + REFRESH + 1 <= 65536 ? 2 : 4;
	const char *gb_env = getenv("PPHPC_GRASS_BYTES");
	if (gb_env) {
		const int v = atoi(gb_env);
		if ((v == 2 && P.R + REFRESH + 1 <= 65536) || v == 4) s->grass_bytes = v;
	}

	s->kpt = 16;
	const char *kpt_env = getenv("PPHPC_SORT_KPT");
	if (kpt_env && atoi(kpt_env) == 8) s->kpt = 8;
	P.sort_tile = WG * s->kpt;
	P.hist_cols = ceil_div(s->cap, P.sort_tile);
	/* Grids: enough groups to fill every CU several times, never more than the
	 * tiles of a full-capacity population; all loops are grid-strided. */
	const uint32_t cu = (uint32_t)s->num_cu;
	s->grid_move = min_u32(ceil_div(s->cap, WG * M_IPT), 8 * cu);
	s->grid_up = min_u32(P.hist_cols, 8 * cu);
	s->grid_down = min_u32(P.hist_cols, (s->rbits == 11 ? 1 : 2) * cu);
	s->grid_act = min_u32(ceil_div(s->cap, A_TILE), 8 * cu);
	s->grid_emit = min_u32(ceil_div(s->cap, A_TILE), 8 * cu);
	s->grid_cells = min_u32(ceil_div(C, WG), 8 * cu);
	s->grid_agents = min_u32(ceil_div(n0 ? n0 : 1, WG), 8 * cu);

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
	k_init_reset<<<min_u32(ceil_div(s->ring_n + s->hist_words, WG), 2048), WG, 0,
		s->stream>>>(s->ctl, s->ring, s->ring_n, s->hist, s->hist_words);
	prof_end(s, pe);
	prof_begin(s, KI_INIT_CELLS, pe);
	if (s->grass_bytes == 1)
		k_init_cells<uint8_t><<<s->grid_cells, WG, 0, s->stream>>>(
			(uint8_t *)s->grass_at, s->ring, s->ctl, s->P);
	else if (s->grass_bytes == 2)
		k_init_cells<uint16_t><<<s->grid_cells, WG, 0, s->stream>>>(
			(uint16_t *)s->grass_at, s->ring, s->ctl, s->P);
	else
		k_init_cells<uint32_t><<<s->grid_cells, WG, 0, s->stream>>>(
			(uint32_t *)s->grass_at, s->ring, s->ctl, s->P);
	prof_end(s, pe);
	prof_begin(s, KI_INIT_AGENTS, pe);
	k_init_agents<<<s->grid_agents, WG, 0, s->stream>>>(s->cell, s->energy,
		s->type, s->ctl, s->P);
	prof_end(s, pe);
	prof_begin(s, KI_INIT_DONE, pe);
	k_init_done<<<1, 64, 0, s->stream>>>(s->ctl, s->stats_dev, s->P);
	prof_end(s, pe);
	prof_begin(s, KI_MOVE0, pe);
	k_move0<<<s->grid_move, WG, 0, s->stream>>>(s->cell, s->energy, s->type,
		s->keys[0], s->vals[0], s->hist, s->ctl, s->P);
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
		/* Countdown at state st from the regrow step a stored modulo 2^w:
		 * d = (a - st) mod 2^w is the countdown if 1 <= d <= R, else 0. */
		const uint32_t st = s->ctl_host->step, R = s->P.R, C = s->P.C;
		const size_t gb = s->grass_bytes;
		std::vector<uint8_t> raw((size_t)C * gb);
		HIP_OK(hipMemcpy(raw.data(), s->grass_at, raw.size(),
			hipMemcpyDeviceToHost));
		const uint32_t mask = gb == 4 ? 0xFFFFFFFFu : ((1u << (8 * gb)) - 1u);
		for (uint32_t c = 0; c < C; c++) {
			uint32_t a;
			if (gb == 1) a = raw[c];
			else if (gb == 2) a = reinterpret_cast<const uint16_t *>(raw.data())[c];
			else a = reinterpret_cast<const uint32_t *>(raw.data())[c];
			const uint32_t d = (a - st) & mask;
			grass[c] = (d >= 1 && d <= R) ? d : 0;
		}
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
	if (sort_tile) *sort_tile = s->P.sort_tile;
	if (action_tile) *action_tile = A_TILE;
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
