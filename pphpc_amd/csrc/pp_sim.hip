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
 *                                            O(1) ring update at the end of
 *                                            tick (k_act_scan, or the emit's
 *                                            group 0 for one replica)
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
 *                                            shuffled order; output offsets
 *                                            (k_act_scan, or each emit group
 *                                            for one replica); k_emit_move:
 *                                            compaction and spawning
 *   reduce_grass1/2, reduce_agent1/2         fused reductions (k_act, then
 *   (pp_gpu.cl:485-736)                      k_act_scan or end_tick_solo) with
 *                                            exact 64-bit sums
 *   stats readback every iteration           device rings of stats rows, one
 *   (pp_gpu.c:766, :840)                     copy per chunk of iterations
 *
 * Replicas. A batch holds independent simulations (replicas, each with its
 * own grid, parameters, seed and iteration count) in one set of buffers:
 * replica r owns the cells [cb_r, cb_r + C_r) of a global cell space, so one
 * sort orders every replica's agents at once and every kernel serves all
 * replicas in the same launch, finding a cell's replica in a table sorted by
 * first cell. Random draws use replica-local agent ids, cells and iteration
 * numbers, so a replica evolves exactly as it would alone. The parameters the
 * reference compiles in as -D constants (pp_gpu.c:1522-1574) are per-replica
 * table entries here. A single simulation (pp_sim_*) is a batch of one.
 *
 * Time. The batch counts iterations in ticks. Replica r, admitted at tick
 * t0_r, is at its local iteration s = tick - t0_r. Grass regrow steps are
 * stored in ticks, so the grass of every replica uses one clock.
 *
 * No kernel waits on another work-group: every cross-group dependency is a
 * kernel boundary, so the step has no spin loops.
 *
 * All arithmetic is integer. The state, its ordering and every random draw are
 * defined so that a single-threaded CPU implementation reproduces the same
 * statistics bit for bit; see the header of each kernel.
 *
 * Device state (capacity = A agents, C cells over all replicas, T sort tiles):
 *   agents (SoA)      cell u32[A], energy i32[A], type u8[A]       9 B/agent
 *   sort pairs        key u32[A], val u32[A], three buffers       24 B/agent
 *   grass             regrow tick modulo 2^8 (R <= 191), 2^16 or 2^32,
 *                     1, 2 or 4 B/cell
 *   grass rings       u64[R_max + 1] per replica slot
 *   tile histograms   u32[radix][T], digit totals u32[radix]
 *   action tiles      u32 output count per 1024 sorted positions
 *   replica tables    parameters, control words, statistics rings
 *   control block     batch counters, tick, error
 */
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
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
constexpr uint32_t NO_REP = 0xFFFFFFFFu;
constexpr uint32_t STATS_RING = 4096; /* stats rows kept per plain simulation */
constexpr uint32_t ITERS_UNLIMITED = 0xFFFFFFFFu; /* iterations of a pp_sim */
constexpr int A_IPT = 4;            /* sorted positions per thread (action) */
constexpr int A_TILE = WG * A_IPT;
static_assert(A_IPT == 2 || A_IPT == 4, "the emit loads a thread's pairs as one vector");
constexpr int A_STAGE = 2 * A_TILE; /* staged outputs per action tile */
constexpr int MIX_SLOTS = 16;       /* replicas per action tile summed in LDS */
constexpr uint32_t K_REG = 5;       /* agents of a cell acted on in registers */
/* ... in the second, rarely taken register path: at 1600 x 1600 the six to
 * nine tiles an iteration that hold a cell of six or seven agents took 12-21 us
 * in the general path against 8.7 for the others, and set the kernel's end.
 * Six: 45.4 -> 43.5 us an iteration (12800 x 12800: 645 -> 640). Seven or
 * eight take more registers than six waves leave (scratch spills): 0.1 us
 * more gained at 1600 x 1600, the 5 us at 12800 x 12800 lost. */
constexpr uint32_t K_REG2 = 6;
template <uint32_t N> struct KConst { static constexpr uint32_t value = N; };
/* Launch shape defaults (pp_batch_set_launch; the tuner tries the others):
 * the action and emit kernels take four times the groups the GPU holds at
 * once, so that groups that finish early are replaced, which evens out tiles
 * of uneven work (measured at 12800 x 12800: 7% faster than one round); the
 * bucket sort takes one round of resident groups (eighths of a round). */
constexpr uint32_t TILE_ROUNDS = 4;
constexpr uint32_t BUCKET_EIGHTHS = 8;
/* Action and emit groups of 512 threads (two positions each) when a tick's
 * tiles fit in one round of such groups: a tile's serial parts (the emit's
 * output rounds, the act's single-agent cells) are then half as long. One
 * 1600 x 1600 replica: 46.4 -> 45.1 us an iteration for the emit, 42.4 ->
 * 41.7 for the act. From 2000 x 2000 to 3200 x 3200, at 12800 x 12800 (8% for
 * the emit) and for several replicas 256 threads are faster; 4000 x 4000 is
 * the exception measured (wide 4% faster). */
constexpr int WIDE_WG = 2 * WG;
constexpr int SCAN_WG = 1024;       /* threads of the one-group scans */
constexpr int SCAN_UNROLL = 8;
/* Grass regrow ticks are stored modulo 2^(8 sizeof(GT)); alive cells are
 * re-stamped once every REFRESH ticks (1/REFRESH of the cells per tick), so
 * every stored value is within R + REFRESH of the tick. */
constexpr uint32_t REFRESH = 64;

/* Packed agent value carried through the sort and the action kernels. An
 * agent whose energy after its action exceeds V_ENERGY fails its replica with
 * PP_ENERGY_OVERFLOW at that iteration (the oracle applies the same limit);
 * energy never grows in the move, so the move needs no check. */
constexpr uint32_t V_ENERGY = 0x00FFFFFFu; /* energy, 24 bits */
constexpr uint32_t V_WOLF = 1u << 24;      /* type (0 sheep, 1 wolf) */
constexpr uint32_t V_REPRO = 1u << 25;     /* reproduction draw < probability */
constexpr uint32_t V_KILLED = 1u << 26;    /* eaten by a wolf this step */
constexpr uint32_t V_CHILD = 1u << 27;     /* reproduced this step */
constexpr uint32_t V_ODD = 1u << 28;       /* energy before the split was odd */

/* Replica states (the values of PP_REPLICA_* in pphpc.h). */
constexpr uint32_t RS_FREE = 0;
constexpr uint32_t RS_ACTIVE = 1;
constexpr uint32_t RS_DONE = 2;
constexpr uint32_t RS_FAILED = 3;

/* Counters k_act accumulates per replica for the statistics row. Two hold a
 * second count in their high 32 bits (each half stays below 2^29, the agent
 * capacity, so no carry crosses): AC_OCC the occupied cells (low) and the
 * agents above the energy limit (high), AC_OUT the outputs (low) and the
 * newborns among them (high). */
enum { AC_SHEEP = 0, AC_WOLVES, AC_SHEEP_EN, AC_WOLVES_EN, AC_EATEN, AC_OCC,
	AC_OUT, NACC };
constexpr unsigned long long HI32 = 1ull << 32;
__host__ __device__ __forceinline__ uint32_t lo32(unsigned long long v) {
	return (uint32_t)v;
}
__host__ __device__ __forceinline__ uint32_t hi32(unsigned long long v) {
	return (uint32_t)(v >> 32);
}

/* Per-replica constants, written by the host at admission. */
struct Rep {
	uint32_t cb;               /* first global cell */
	uint32_t C, gx, gy, R;
	uint32_t sheep_gain, wolves_gain;
	uint32_t sheep_thr, wolves_thr;
	uint32_t sheep_prob, wolves_prob;
	uint32_t init_sheep, n0;
	uint32_t key0, key1;
	uint32_t t0;               /* tick of admission */
	uint32_t iters;            /* done after this many iterations */
	uint32_t cap;              /* agents allowed after births */
	uint32_t stats_rows;       /* length of its stats ring */
	uint32_t gx_magic, gx_shift; /* division by gx: see div_by */
	uint32_t pad;
	unsigned long long *ring;  /* R + 1 counters of dead cells per regrow tick */
	pp_stats64_t *stats;       /* stats ring */
};

/* Per-replica control words, updated by the kernels. */
struct RepCtl {
	uint32_t n;        /* agents of its current state */
	uint32_t base;     /* position of its first agent in the batch state */
	uint32_t step;     /* completed iterations */
	uint32_t status;   /* RS_* */
	int32_t error;     /* PP_* code of a failed replica */
	uint32_t err_step; /* iteration at which it failed */
	uint32_t pad[2];
	unsigned long long dead_cells;  /* grass cells with countdown > 0 */
	unsigned long long cd_sum;      /* sum of grass countdowns */
	unsigned long long acc[NACC];   /* this iteration, AC_* */
	unsigned long long agent_steps; /* agents at the start, summed */
};

struct Ctl {
	uint32_t n;        /* positions of the batch state */
	uint32_t live[2];  /* survivors of the move, indexed by tick parity */
	uint32_t tick;     /* iterations run by this batch */
	int32_t error;     /* 0 or a PP_* code; every kernel is a no-op once set */
	uint32_t nrep;     /* replicas in the lookup table */
	uint32_t nbuckets; /* buckets of the replicas in the table */
	uint32_t nmixed;   /* action tiles listed for k_act_mixed this tick */
	uint32_t err_step; /* tick at which error was set */
	int32_t err_rep;   /* slot of the replica involved, -1 for the batch */
	/* the emit that runs with this tick (the last of a run of ticks) writes
	 * the SoA state; only the host reads it */
	uint32_t state_tick;
	/* the tick the emit moves for, published by the action kernel when the
	 * table holds one replica: the emit's work-group 0 then ends the tick
	 * (advances Ctl::tick) while the others still run, so they read this */
	uint32_t emit_tick;
	/* sums over completed ticks of: agents at the start, alive after the
	 * move, after actions, occupied cells, grass eaten, births */
	unsigned long long agent_steps, sum_live, sum_out, sum_occ, sum_eaten,
		sum_births;
};

/* The first eight words of Ctl, which the step kernels start from: read
 * with one load (a chain of dependent reads - the error flag, then the
 * tick, then the live count of its parity - costs a round trip each). */
struct CtlHead {
	uint32_t n, live0, live1, tick;
	int32_t error;
	uint32_t nrep, nbuckets, nmixed;
	__device__ __forceinline__ uint32_t live(uint32_t t) const {
		return (t & 1u) ? live1 : live0;
	}
};
static_assert(offsetof(Ctl, nmixed) == 28, "CtlHead covers Ctl's first 8 words");

__device__ __forceinline__ CtlHead ctl_head(const Ctl *ctl) {
	typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));
	const u32x8 v = *reinterpret_cast<const u32x8 *>(ctl);
	CtlHead h;
	h.n = v[0];
	h.live0 = v[1];
	h.live1 = v[2];
	h.tick = v[3];
	h.error = (int32_t)v[4];
	h.nrep = v[5];
	h.nbuckets = v[6];
	h.nmixed = v[7];
	return h;
}

/* Batch-wide launch constants. */
struct Params {
	uint32_t cap;        /* agent capacity of the buffers */
	uint32_t passes;     /* radix passes over the global cell index */
	uint32_t digit_bits; /* bits per pass (<= 11) */
	uint32_t sort_tile;  /* positions per sort tile */
	uint32_t hist_cols;  /* tile capacity of the histogram table */
	uint32_t C;          /* cells of the global cell space */
	uint32_t R;          /* largest grass restart allowed */
	uint32_t bshift;     /* bucket blocks are 2^bshift cells of a grid row */
	uint32_t bucket_on;  /* sort by row buckets; the emit counts for them */
	uint32_t nb_cap;     /* bucket capacity of the count tables */
	uint32_t cnt_top;    /* highest level of the bucket count tables in use */
	uint32_t one_slot;   /* slot of the replica when the table holds one */
	uint32_t solo;       /* the table holds one replica: no k_act_scan */
	uint32_t act_l1;     /* offset of the per-64-tile sums in the action counts */
	unsigned long long *part;  /* partial-sum slots (PART_*) */
};

/* Partial sums that every work-group of a kernel adds to: PART_SLOTS slots
 * 128 bytes apart, a group adding to slot (group index mod PART_SLOTS).
 * Atomic adds of many work-groups to one address serialise (about 11 ns each
 * on MI355X: 600 groups adding to one counter lengthen a kernel by about
 * 6 us); spread over 64 slots they cost nothing measurable. Words
 * [0, NACC) of a slot: the action counters of a one-replica batch, summed by
 * k_act_scan; word PART_LIVE: the move's survivors, summed into Ctl::live by
 * the next tick's first kernel. Both consumers clear what they sum. */
constexpr uint32_t PART_SLOTS = 64, PART_STRIDE = 16, PART_LIVE = 8;
static_assert(PART_SLOTS == (uint32_t)WAVE, "one wave sums the slots");
static_assert(NACC <= (int)PART_LIVE && PART_LIVE < PART_STRIDE,
	"slot layout");
/* The action counters of a one-replica batch go to words [0, NACC) of a
 * slot at even ticks and [PART_ODD, PART_ODD + NACC) at odd ones: the emit
 * reads tick t's while the next tick's first kernel clears them (fold_live),
 * before tick t + 2's actions add to them again. */
constexpr uint32_t PART_ODD = PART_LIVE + 1;
static_assert(PART_ODD + NACC <= PART_STRIDE, "slot layout");
__host__ __device__ __forceinline__ uint32_t part_acc(uint32_t tick) {
	return (tick & 1u) ? PART_ODD : 0u;
}

/* Replica tables in device memory: constants and control words by slot, and
 * the replicas of the batch sorted by first cell (cbs ascending, order[k] =
 * slot of the k-th). */
struct Tables {
	const Rep *rep;
	RepCtl *rctl;
	const uint32_t *cbs;
	const uint32_t *order;
	const uint32_t *bbase;  /* first bucket of the k-th replica; [nrep] = all */
	/* per chunk of 2^CREP_SHIFT cells: the last table entry whose first cell
	 * is at or before the chunk's first cell (0 before the first) */
	const uint32_t *crep;
};
constexpr uint32_t CREP_SHIFT = 12;

/* ---------------------------------------------------------------- helpers */

/* LDS accesses through explicitly LDS-typed pointers. Where a branch picks
 * between an LDS array and a global one, plain pointers let the compiler
 * merge both sides into one flat (generic) access, and every flat access
 * waits for all outstanding global loads and stores (vmcnt(0) and
 * lgkmcnt(0)), which stalls prefetches and store streams behind it. */
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
typedef __attribute__((address_space(1))) unsigned long long glb_u64;

/* p[i] through a 32-bit byte offset: agent capacities stay below 2^29, so
 * the offsets of agent arrays fit 32 bits, and the compiler can use the
 * scalar base + 32-bit offset form of a global access instead of 64-bit
 * address arithmetic in every lane. */
template <typename T>
__device__ __forceinline__ T &at32(T *p, uint32_t i) {
	return *reinterpret_cast<T *>(reinterpret_cast<char *>(p) +
		(uint32_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const T &at32(const T *p, uint32_t i) {
	return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) +
		(uint32_t)(i * (uint32_t)sizeof(T)));
}

/* Grass of global cell c: one-byte grass through a 32-bit byte offset (any
 * cell index fits), wider grass by plain indexing. */
template <typename GT>
__device__ __forceinline__ GT &grass_cell(GT *g, uint32_t c) {
	if constexpr (sizeof(GT) == 1) return at32(g, c);
	else return g[c];
}

/* Waits until every global load and store this wave issued has completed
 * (s_waitcnt vmcnt(0)). Global accesses retire in issue order, so where
 * loads are issued conditionally, or stores follow them, the compiler can
 * only wait on a load with vmcnt(0), which then also waits for stores just
 * issued. One explicit wait at a point where the loads have had time to
 * arrive, before any store, leaves the compiler no load to wait for
 * afterwards. */
__device__ __forceinline__ void wait_global() {
	__builtin_amdgcn_s_waitcnt(0x0F70);   /* vmcnt(0), expcnt/lgkmcnt free */
}

__device__ __forceinline__ uint32_t lds_ld(const uint32_t *p) {
	return *(const lds_u32 *)p;
}
__device__ __forceinline__ void lds_st(uint32_t *p, uint32_t v) {
	*(lds_u32 *)p = v;
}
__device__ __forceinline__ void lds_add(uint32_t *p, uint32_t v) {
	__hip_atomic_fetch_add((lds_u32 *)p, v, __ATOMIC_RELAXED,
		__HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(unsigned long long *p,
	unsigned long long v) {
	__hip_atomic_fetch_add((lds_u64 *)p, v, __ATOMIC_RELAXED,
		__HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void glb_add(unsigned long long *p,
	unsigned long long v) {
	__hip_atomic_fetch_add((glb_u64 *)p, v, __ATOMIC_RELAXED,
		__HIP_MEMORY_SCOPE_AGENT);
}

/* Inclusive prefix sum of a 32-bit value over the lanes of the wave (all
 * lanes active), with DPP: row shifts by 1, 2, 4 and 8 scan each row of 16
 * lanes, then row broadcasts of lanes 15 and 31 carry the rows' totals. */
__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t v) {
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false); /* row_shr:1 */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false); /* row_shr:2 */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false); /* row_shr:4 */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false); /* row_shr:8 */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); /* row_bcast:15 */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); /* row_bcast:31 */
	return v;
}

/* The value of the previous lane (lane 0: its own), with DPP wave_shr:1 -
 * a register move across lanes, no LDS round trip. */
__device__ __forceinline__ uint32_t prev_lane(uint32_t v) {
	return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x138, 0xf, 0xf,
		false);
}

/* Work-group exclusive scan of one value per thread (NT threads). */
template <int NT>
__device__ __forceinline__ uint32_t wg_excl_scan(uint32_t v, uint32_t *s_tmp,
	uint32_t &total) {
	const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
	const uint32_t x = wave_incl_scan32(v);
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

/* Adds the survivors of the last move, which the emit's work-groups left in
 * the partial-sum slots, to live[tick & 1], and clears the slots; clears the
 * action counters of the previous tick (read by its emit, reused by the
 * next tick's actions). Called by the first wave of one work-group of a
 * tick's first kernel (no other work-group writes Ctl::live then, and the
 * kernels that read it follow). */
__device__ __forceinline__ void fold_live(Ctl *ctl, unsigned long long *part,
	uint32_t tick) {
	const int lane = threadIdx.x;
	unsigned long long &p = part[lane * PART_STRIDE + PART_LIVE];
	const unsigned long long s = wave_sum(p);
	p = 0;      /* issued after the load returned: s depends on it */
	unsigned long long *q = part + lane * PART_STRIDE + part_acc(tick + 1);
#pragma unroll
	for (int i = 0; i < NACC; i++) q[i] = 0;
	if (lane == 0 && s) ctl->live[tick & 1] += (uint32_t)s;
}

/* Sum of a 32-bit value over the 64 lanes of the wave (all lanes active),
 * returned in every lane: DPP quad permutes, row shifts and row broadcasts
 * (no LDS round trips, unlike lane shuffles). Lane 63 ends with the total. */
__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false);  /* quad [1,0,3,2] */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false);  /* quad [2,3,0,1] */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false); /* row_shr:4 */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false); /* row_shr:8 */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); /* row_bcast:15 */
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); /* row_bcast:31 */
	return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

/* Work-group sums of K counters (NT threads), added with one atomic each to
 * dst[k]; v is cleared. */
template <int K, int NT = WG>
__device__ __forceinline__ void wg_add(unsigned long long (&v)[K],
	unsigned long long (*s_red)[K], unsigned long long *dst) {
	const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
	for (int k = 0; k < K; k++) {
		v[k] = wave_sum(v[k]);
		if (lane == 0) s_red[w][k] = v[k];
		v[k] = 0;
	}
	__syncthreads();
	if (threadIdx.x < K) {
		unsigned long long t = 0;
		for (int i = 0; i < NT / WAVE; i++) t += s_red[i][threadIdx.x];
		if (t) atomicAdd(&dst[threadIdx.x], t);
	}
	__syncthreads();
}

__device__ __forceinline__ void set_error(Ctl *ctl, int32_t code,
	uint32_t tick, int32_t rep) {
	if (atomicCAS(&ctl->error, 0, code) == 0) {
		ctl->err_step = tick;
		ctl->err_rep = rep;
	}
}

/* Table index of the replica owning global cell c (a cell of some replica):
 * the largest k with cbs[k] <= c. */
__device__ __forceinline__ uint32_t rep_index(const Tables &T, uint32_t nrep,
	uint32_t c) {
	if (nrep <= 1) return 0;
	/* the chunk's entry, then past the replicas that start inside the
	 * chunk before c (none where replicas span more than a chunk) */
	uint32_t k = T.crep[c >> CREP_SHIFT];
	while (k + 1 < nrep && T.cbs[k + 1] <= c) k++;
	return k;
}

/* Grass of a cell, stored as its regrow tick a (the cell is alive at tick t
 * iff a <= t; its countdown is max(0, a - t)) modulo 2^bits of GT. Exact for
 * every replica with R <= Rmax: alive cells are within REFRESH ticks of their
 * stored value, dead ones within R. */
template <typename GT>
__device__ __forceinline__ bool grass_alive_at(GT g, uint32_t t, uint32_t Rmax) {
	/* dead iff a in [t + 1, t + R], i.e. (t - a) mod 2^w >= 2^w - R */
	const uint32_t w = 8u * sizeof(GT);
	const uint32_t delta = (t - (uint32_t)g) & (w == 32 ? 0xFFFFFFFFu
		: ((1u << w) - 1u));
	const uint32_t dead_from = (w == 32 ? 0u : (1u << w)) - Rmax;
	return delta < dead_from;
}

/* Adds 1 to s_hist[d] for every lane of the wave, merging runs of equal
 * digits in consecutive lanes first (agents arrive ordered by cell, so high
 * digits form long runs). NO_DIGIT lanes count nothing. */
__device__ __forceinline__ void wave_hist_add(uint32_t *s_hist, uint32_t d) {
	const int lane = threadIdx.x & (WAVE - 1);
	const uint32_t prev = prev_lane(d);
	const bool head = lane == 0 || d != prev;
	const uint64_t heads = __ballot(head);
	if (head && d != NO_DIGIT) {
		const uint64_t after = heads & ~((2ull << lane) - 1ull);
		const uint32_t next = after ? (uint32_t)__builtin_ctzll(after)
			: (uint32_t)WAVE;
		atomicAdd(&s_hist[d], next - (uint32_t)lane);
	}
}

/* What the move needs of a replica. */
struct RM {
	uint32_t cb, C, gx, gy, sheep_prob, wolves_prob, key0, key1;
	uint32_t gxm, gxs;  /* division by gx (div_by) */
	uint32_t step;   /* local iteration the moved agents act in */
	uint32_t base;   /* batch position of its local agent id 0 */
	uint32_t bbase;  /* its first bucket */
	uint32_t nb;     /* bucket blocks per grid row */
	uint32_t h;      /* grid rows per bucket */
	bool active;
};

/* Blocks of 2^S cells per grid row of width gx. */
__host__ __device__ __forceinline__ uint32_t row_blocks(uint32_t gx, uint32_t S) {
	return (gx + (1u << S) - 1u) >> S;
}

/* Grid rows per bucket: rows no wider than a block share a bucket, as many
 * whole rows as fit in 2^S cells; a wider row is cut into blocks. */
__host__ __device__ __forceinline__ uint32_t rows_per_bucket(uint32_t gx,
	uint32_t S) {
	return gx > (1u << S) ? 1u : (1u << S) / gx;
}

/* Division of any 32-bit n by a fixed d with one high multiply, a subtract
 * and two shifts (the branch-free unsigned method of the libdivide library:
 * l = ceil(log2 d), magic = floor(2^32 (2^l - d) / d) + 1, shifts min(l, 1)
 * and max(l - 1, 0) packed as shift | shift2 << 8). Exact for all n. */
__host__ __device__ __forceinline__ void div_magic(uint32_t d, uint32_t *magic,
	uint32_t *shift) {
	uint32_t l = 0;
	while (l < 32 && (1ull << l) < d) l++;
	*magic = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
	*shift = (l ? 1u : 0u) | ((l ? l - 1 : 0u) << 8);
}

__device__ __forceinline__ uint32_t div_by(uint32_t n, uint32_t magic,
	uint32_t shift) {
	const uint32_t t = __umulhi(magic, n);
	return (t + ((n - t) >> (shift & 0xFFu))) >> (shift >> 8);
}

/* Bucket of local cell (x, y). */
__device__ __forceinline__ uint32_t bucket_of(uint32_t bbase, uint32_t nb,
	uint32_t h, uint32_t x, uint32_t y, uint32_t S) {
	return bbase + (h == 1 ? y : y / h) * nb + (x >> S);
}

/* The move constants of the k-th replica of the table at tick `tick`. */
__device__ __forceinline__ RM load_rm(const Tables &T, uint32_t k,
	uint32_t tick, uint32_t S, uint32_t slot_hint = NO_REP) {
	const uint32_t slot = slot_hint != NO_REP ? slot_hint : T.order[k];
	const Rep &rp = T.rep[slot];
	const RepCtl &rc = T.rctl[slot];
	RM m;
	m.cb = rp.cb;
	m.C = rp.C;
	m.gx = rp.gx;
	m.gy = rp.gy;
	m.gxm = rp.gx_magic;
	m.gxs = rp.gx_shift;
	m.bbase = T.bbase[k];
	m.nb = row_blocks(rp.gx, S);
	m.h = rows_per_bucket(rp.gx, S);
	m.sheep_prob = rp.sheep_prob;
	m.wolves_prob = rp.wolves_prob;
	m.key0 = rp.key0;
	m.key1 = rp.key1;
	m.step = tick - rp.t0;
	m.base = rc.base;
	m.active = rc.status == RS_ACTIVE;
	return m;
}

/* load_rm for a replica index uniform over the wave, its fields moved to
 * scalar registers (later uses then wait on no vector load). */
__device__ __forceinline__ RM load_rm_uniform(const Tables &T, uint32_t k,
	uint32_t tick, uint32_t S, uint32_t slot_hint = NO_REP) {
	RM m = load_rm(T, k, tick, S, slot_hint);
	m.cb = __builtin_amdgcn_readfirstlane(m.cb);
	m.C = __builtin_amdgcn_readfirstlane(m.C);
	m.gx = __builtin_amdgcn_readfirstlane(m.gx);
	m.gy = __builtin_amdgcn_readfirstlane(m.gy);
	m.gxm = __builtin_amdgcn_readfirstlane(m.gxm);
	m.gxs = __builtin_amdgcn_readfirstlane(m.gxs);
	m.sheep_prob = __builtin_amdgcn_readfirstlane(m.sheep_prob);
	m.wolves_prob = __builtin_amdgcn_readfirstlane(m.wolves_prob);
	m.key0 = __builtin_amdgcn_readfirstlane(m.key0);
	m.key1 = __builtin_amdgcn_readfirstlane(m.key1);
	m.step = __builtin_amdgcn_readfirstlane(m.step);
	m.base = __builtin_amdgcn_readfirstlane(m.base);
	m.bbase = __builtin_amdgcn_readfirstlane(m.bbase);
	m.nb = __builtin_amdgcn_readfirstlane(m.nb);
	m.h = __builtin_amdgcn_readfirstlane(m.h);
	m.active = __builtin_amdgcn_readfirstlane((uint32_t)m.active) != 0;
	return m;
}

/* One move of a replica's agent with local id `id` at local cell (x, y):
 * Philox block (id, step, MOVE, 0) under the replica's key; direction U[0,5)
 * from word 0 (0 stay, 1 +x, 2 -x, 3 +y, 4 -y, torus wrap: the table of
 * move_agent, pp_gpu.cl:780-803); reproduction draw U[0,100) from word 1,
 * kept as one bit (draw < probability of the agent's type) for the action
 * phase. Energy -1; the agent dies below 1 (pphpc.nlogo:66-70,
 * pp_gpu.cl:806-815): key DEAD_KEY, dropped by the sort. Returns the global
 * key (cb + new local cell); *val gets the packed value and *bkt the bucket
 * of the new cell (blocks of 2^S cells per row), NO_DIGIT for the dead. */
__device__ __forceinline__ uint32_t move_agent(uint32_t id, uint32_t x,
	uint32_t y, int32_t energy, uint32_t wolf, const RM &m, uint32_t S,
	uint32_t *val, uint32_t *bkt) {
	const uint4 r = pp_philox(id, m.step, PP_STREAM_MOVE, 0, m.key0, m.key1);
	const uint32_t dir = pp_bounded(r.x, 5);
	const uint32_t dice = pp_bounded(r.y, 100);
	const int32_t e = energy - 1;
	*val = 0;
	*bkt = NO_DIGIT;
	if (e < 1) return DEAD_KEY;
	uint32_t nx = x, ny = y;
	if (dir == 1) nx = (x == m.gx - 1) ? 0u : x + 1;
	else if (dir == 2) nx = (x == 0) ? m.gx - 1 : x - 1;
	else if (dir == 3) ny = (y == m.gy - 1) ? 0u : y + 1;
	else if (dir == 4) ny = (y == 0) ? m.gy - 1 : y - 1;
	const uint32_t prob = wolf ? m.wolves_prob : m.sheep_prob;
	*val = ((uint32_t)e & V_ENERGY) | (wolf ? V_WOLF : 0u)
		| (dice < prob ? V_REPRO : 0u);
	*bkt = bucket_of(m.bbase, m.nb, m.h, nx, ny, S);
	return m.cb + ny * m.gx + nx;
}

/* Bucket count tables: three levels of 64-ary sums, so that any prefix sum
 * is at most 3 x 63 reads that one wave does at once. Level k holds, for
 * every q, the count of buckets [64^k q, 64^k (q + 1)); it has
 * (nb_cap >> 6k) + 1 entries and starts at cnt_off(nb_cap, k). Layouts of
 * up to 2^18 buckets fit (more sort with the radix sort). Only levels
 * 0 .. cnt_top are kept, cnt_top the highest with an entry above 0 for the
 * layout (NB > 64^cnt_top). */
constexpr int CNT_LEVELS = 3;
constexpr uint32_t CNT_MAX_BUCKETS = 1u << 18;

__host__ __device__ __forceinline__ uint32_t cnt_off(uint32_t nbc, int lvl) {
	uint32_t o = 0;
	for (int k = 0; k < lvl; k++) o += (nbc >> (6 * k)) + 1;
	return o;
}

/* After the levels of the old-cell table: per bucket, the moved agents on
 * the first cell of its block and on the last (row blocks of grids wider
 * than a block), where the agents of the neighbouring blocks' edge cells
 * start and end. */
__host__ __device__ __forceinline__ uint32_t edge_first_off(uint32_t nbc) {
	return cnt_off(nbc, CNT_LEVELS);
}
__host__ __device__ __forceinline__ uint32_t edge_last_off(uint32_t nbc) {
	return cnt_off(nbc, CNT_LEVELS) + nbc + 1;
}
__host__ __device__ __forceinline__ uint32_t cnt_words(uint32_t nbc) {
	return cnt_off(nbc, CNT_LEVELS) + 2 * (nbc + 1);
}

/* LDS windows of the emit's bucket counts: WIN level-0 entries from w0 and
 * L1W level-1 entries from u0 (level 1 covers 64 buckets an entry, so the
 * rows above and below of any grid up to 4096 blocks wide fall in it). */
constexpr uint32_t L1W = 130;
constexpr uint32_t L2W = 4;      /* level-2 entries the level-1 window spans */

/* Adds 1 to the count of bucket `bkt` for every lane of the wave (NO_DIGIT
 * lanes count nothing), merging runs of equal buckets in consecutive lanes
 * first: level 0 into the LDS window s_win of WIN buckets from w0 when it
 * falls there, else straight into the global table; levels 1 .. top into
 * the LDS window s_l1 from u0, else straight into the global levels. */
template <uint32_t WIN>
__device__ __forceinline__ void wave_bucket_add(uint32_t *s_win,
	uint32_t *s_l1, uint32_t *g_cnt, uint32_t nbc, uint32_t top, uint32_t w0,
	uint32_t u0, uint32_t bkt) {
	const int lane = threadIdx.x & (WAVE - 1);
	const uint32_t prev = prev_lane(bkt);
	const bool head = lane == 0 || bkt != prev;
	const uint64_t heads = __ballot(head);
	if (head && bkt != NO_DIGIT) {
		const uint64_t after = heads & ~((2ull << lane) - 1ull);
		const uint32_t next = after ? (uint32_t)__builtin_ctzll(after)
			: (uint32_t)WAVE;
		const uint32_t run = next - (uint32_t)lane;
		const uint32_t o = bkt - w0;
		if (bkt >= w0 && o < WIN) lds_add(&s_win[o], run);
		else atomicAdd(&g_cnt[bkt], run);
		if (top >= 1) {
			const uint32_t q1 = bkt >> 6, o1 = q1 - u0;
			if (q1 >= u0 && o1 < L1W) {
				lds_add(&s_l1[o1], run);
			} else {
				for (uint32_t k = 1; k <= top; k++)
					atomicAdd(&g_cnt[cnt_off(nbc, (int)k) + (bkt >> (6 * k))], run);
			}
		}
	}
}

/* ------------------------------------------------------------------- init */

/* Empties the batch: no agents, no replicas, tick 0, the pass-0 histograms
 * and the partial-sum slots zeroed. */
__global__ __launch_bounds__(WG) void k_batch_reset(Ctl *ctl, uint32_t *hist,
	uint32_t hist_words, unsigned long long *part) {
	const uint32_t gtid = blockIdx.x * WG + threadIdx.x;
	for (uint32_t i = gtid; i < hist_words; i += gridDim.x * WG) hist[i] = 0;
	for (uint32_t i = gtid; i < PART_SLOTS * PART_STRIDE; i += gridDim.x * WG)
		part[i] = 0;
	if (gtid == 0) memset(ctl, 0, sizeof(Ctl));
}

/* Marks the last of the next `ticks` ticks: its emit writes the SoA state
 * (the other emits skip it; no kernel reads it, the host does after a run). */
__global__ void k_state_tick(Ctl *ctl, uint32_t ticks) {
	if (threadIdx.x == 0 && blockIdx.x == 0 && !ctl->error)
		ctl->state_tick = ctl->tick + ticks;
}

/* Clears one replica slot's control words and grass ring. */
__global__ __launch_bounds__(WG) void k_rep_reset(Tables T, uint32_t slot) {
	const Rep &rp = T.rep[slot];
	for (uint32_t i = threadIdx.x; i <= rp.R; i += WG) rp.ring[i] = 0;
	if (threadIdx.x == 0) memset(&T.rctl[slot], 0, sizeof(RepCtl));
}

/* Cells of one replica: grass alive with probability 1/2 (draw word 0), else
 * countdown 1 + U[0, R) (word 1), Philox counter (local cell, 0, INIT_CELL,
 * 0). Follows init_cell (pp_gpu.cl:368-398) and pphpc.nlogo:27-35. A cell with
 * countdown g at local iteration 0 is stored as regrow tick t0 + g; dead
 * cells are counted in ring slot (t0 + g) mod (R + 1). */
template <typename GT>
__global__ __launch_bounds__(WG) void k_init_cells(GT *grass_at, Tables T,
	uint32_t slot) {
	__shared__ uint32_t s_hist[4096];
	__shared__ unsigned long long s_red[NWAVE][2];
	const Rep rp = T.rep[slot];
	const uint32_t M = rp.R + 1;
	const bool lds_hist = M <= 4096;
	if (lds_hist)
		for (uint32_t i = threadIdx.x; i < M; i += WG) s_hist[i] = 0;
	__syncthreads();
	unsigned long long v[2] = {0, 0};
	for (uint32_t c = blockIdx.x * WG + threadIdx.x; c < rp.C;
		c += gridDim.x * WG) {
		const uint4 r = pp_philox(c, 0, PP_STREAM_INIT_CELL, 0, rp.key0, rp.key1);
		uint32_t g = 0;
		if (pp_bounded(r.x, 2) != 1) g = 1 + pp_bounded(r.y, rp.R);
		grass_at[rp.cb + c] = (GT)(rp.t0 + g);
		if (g) {
			v[0] += 1;
			v[1] += g;
			const uint32_t k = (rp.t0 + g) % M;
			if (lds_hist) atomicAdd(&s_hist[k], 1u);
			else atomicAdd(&rp.ring[k], 1ull);
		}
	}
	wg_add<2>(v, s_red, &T.rctl[slot].dead_cells);
	if (lds_hist)
		for (uint32_t i = threadIdx.x; i < M; i += WG)
			if (s_hist[i]) atomicAdd(&rp.ring[i], (unsigned long long)s_hist[i]);
}

/* Agents of one replica at batch positions [pos, pos + n0): local ids
 * 0..init_sheep-1 are sheep, the rest wolves; x = U[0,GX) (word 0),
 * y = U[0,GY) (word 1), energy 1 + U[0, 2*gain) (word 2), Philox counter
 * (local id, 0, INIT_AGENT, 0). Follows init_agent (pp_gpu.cl:406-439) and
 * pphpc.nlogo:38-53. */
__global__ __launch_bounds__(WG) void k_init_agents(uint32_t *cell,
	int32_t *energy, uint8_t *type, Tables T, uint32_t slot, uint32_t pos) {
	__shared__ unsigned long long s_red[NWAVE][4];
	const Rep rp = T.rep[slot];
	unsigned long long v[4] = {0, 0, 0, 0};
	for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < rp.n0;
		i += gridDim.x * WG) {
		const uint4 r = pp_philox(i, 0, PP_STREAM_INIT_AGENT, 0, rp.key0, rp.key1);
		const bool wolf = i >= rp.init_sheep;
		const uint32_t gain = wolf ? rp.wolves_gain : rp.sheep_gain;
		const uint32_t x = pp_bounded(r.x, rp.gx);
		const uint32_t y = pp_bounded(r.y, rp.gy);
		const int32_t e = 1 + (int32_t)pp_bounded(r.z, 2 * gain);
		cell[pos + i] = rp.cb + y * rp.gx + x;
		energy[pos + i] = e;
		type[pos + i] = wolf ? PP_WOLF : PP_SHEEP;
		if (wolf) { v[1] += 1; v[3] += (unsigned long long)e; }
		else { v[0] += 1; v[2] += (unsigned long long)e; }
	}
	wg_add<4>(v, s_red, &T.rctl[slot].acc[0]);
}

/* One statistics row from the counters acc (AC_*), the live grass cells
 * and the countdown sum. */
__device__ __forceinline__ void write_row(const unsigned long long *acc,
	unsigned long long grass, unsigned long long grass_en, pp_stats64_t *row) {
	row->sheep = acc[AC_SHEEP];
	row->wolves = acc[AC_WOLVES];
	row->grass = grass;
	row->sheep_en = acc[AC_SHEEP_EN];
	row->wolves_en = acc[AC_WOLVES_EN];
	row->grass_en = grass_en;
	row->errors = 0;
}
__device__ void write_row(const unsigned long long *acc, const RepCtl &rc,
	const Rep &rp, pp_stats64_t *row) {
	write_row(acc, (unsigned long long)rp.C - rc.dead_cells, rc.cd_sum, row);
}

/* Row 0 of one replica placed at batch position pos = ctl->n; its agents join
 * the batch state. */
__global__ void k_rep_init_done(Ctl *ctl, Tables T, uint32_t slot,
	uint32_t pos) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	const Rep &rp = T.rep[slot];
	RepCtl &rc = T.rctl[slot];
	write_row(rc.acc, rc, rp, &rp.stats[0]);
	for (int i = 0; i < NACC; i++) rc.acc[i] = 0;
	rc.n = rp.iters ? rp.n0 : 0;
	rc.base = pos;
	rc.step = 0;
	rc.status = rp.iters ? RS_ACTIVE : RS_DONE;
	ctl->n = pos + rp.n0;
}

/* Move for one replica's freshly initialised agents at batch positions
 * [pos, pos + n0) (SoA in local id order): their pairs and the live count of
 * this tick. They are appended after the agents of the batch, out of cell
 * order, so the next tick sorts with the general radix sort. */
__global__ __launch_bounds__(WG) void k_move0(const uint32_t *__restrict__ cell,
	const int32_t *__restrict__ energy, const uint8_t *__restrict__ type,
	uint32_t *__restrict__ keys, uint32_t *__restrict__ vals, Ctl *ctl,
	Tables T, uint32_t slot, uint32_t pos, Params P) {
	__shared__ unsigned long long s_red[NWAVE];
	if (ctl->error) return;
	const uint32_t tick = ctl->tick;
	const Rep &rp = T.rep[slot];
	const RepCtl &rc = T.rctl[slot];
	RM m;
	m.cb = rp.cb;
	m.C = rp.C;
	m.gx = rp.gx;
	m.gy = rp.gy;
	m.sheep_prob = rp.sheep_prob;
	m.wolves_prob = rp.wolves_prob;
	m.key0 = rp.key0;
	m.key1 = rp.key1;
	m.step = tick - rp.t0;
	m.base = pos;
	m.bbase = 0;
	m.nb = row_blocks(rp.gx, P.bshift);
	m.h = rows_per_bucket(rp.gx, P.bshift);
	m.active = rc.status == RS_ACTIVE;
	const uint32_t n = rp.n0;
	unsigned long long live = 0;
	for (uint32_t l = blockIdx.x * WG + threadIdx.x; l < n; l += gridDim.x * WG) {
		const uint32_t g = pos + l;
		uint32_t k = DEAD_KEY, v = 0, bkt;
		if (m.active) {
			const uint32_t c = cell[g] - m.cb;
			const uint32_t y = c / m.gx, x = c - y * m.gx;
			k = move_agent(l, x, y, energy[g], type[g], m, P.bshift, &v, &bkt);
		}
		keys[g] = k;
		vals[g] = v;
		live += k != DEAD_KEY;
	}
	const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
	const unsigned long long tsum = wave_sum(live);
	if (lane == 0) s_red[w] = tsum;
	__syncthreads();
	if (tid == 0) {
		unsigned long long sum = 0;
		for (int i = 0; i < NWAVE; i++) sum += s_red[i];
		if (sum) atomicAdd(&ctl->live[tick & 1], (uint32_t)sum);
	}
}

/* ------------------------------------------------------------------- sort */

/* Radix sort of the (cell, value) pairs by cell, least significant digit
 * first, digit_bits per pass. Every pass is stable, so survivors end ordered
 * by (cell, agent id) -- the key (cell << 32 | agent id) -- and the first pass
 * drops dead agents (DEAD_KEY). A pass is three kernels:
 *   k_sort_up    digit histogram of every tile
 *   k_sort_scan  per digit, exclusive prefix over tiles and the digit total
 *   k_sort_down  per tile: rank in LDS, scatter to bucket start + tile prefix
 * Tile t covers positions [t * sort_tile, (t + 1) * sort_tile). The table
 * hist[d * hist_cols + t] holds tile counts, then tile offsets. */

__device__ __forceinline__ uint32_t pass_len(const Ctl *ctl, uint32_t pass) {
	return pass == 0 ? ctl->n : ctl->live[ctl->tick & 1];
}

template <int KPT>
__global__ __launch_bounds__(WG) void k_sort_up(const uint32_t *__restrict__ kin,
	uint32_t *__restrict__ hist, Ctl *ctl, uint32_t pass, Params P) {
	constexpr uint32_t TILE = WG * KPT;
	__shared__ uint32_t s_hist[RADIX_MAX];
	if (ctl->error) return;
	/* the tick's first kernel: the last move's survivors into live, which
	 * the later passes read (pass 0 sorts all n positions) */
	if (pass == 0 && blockIdx.x == 0 && threadIdx.x < WAVE)
		fold_live(ctl, P.part, ctl->tick);
	const uint32_t n = pass_len(ctl, pass);
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
		/* Re-stamp the alive cells of slice (t mod REFRESH) with t, which
		 * keeps their stored regrow tick within the modular window; the
		 * state does not change (alive stays alive, countdown 0). Cells of
		 * no replica are stamped too, harmlessly. */
		const uint32_t s = ctl->tick;
		const uint32_t per = (P.C + REFRESH - 1) / REFRESH;
		const uint32_t c0 = (s % REFRESH) * per;
		const uint32_t c1 = min(c0 + per, P.C);
		for (uint32_t c = c0 + blockIdx.x * WG + threadIdx.x; c < c1;
			c += gridDim.x * WG) {
			const GT g = grass_at[c];
			if (grass_alive_at<GT>(g, s, P.R) && g != (GT)s) grass_at[c] = (GT)s;
		}
	}
	const uint32_t n = pass_len(ctl, pass);
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
	Ctl *ctl, uint32_t pass, Params P) {
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
	const uint32_t n = pass_len(ctl, pass);
	const uint32_t ntiles = (n + TILE - 1) / TILE;
	const uint32_t shift = pass * P.digit_bits;
	const uint32_t radix = 1u << P.digit_bits, dmask = radix - 1u;
	const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
	const uint64_t lt_mask = (1ull << lane) - 1ull;
	bool bad = false;

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
			/* dst < n whenever the histograms describe the input; the guard
			 * turns an inconsistent table into an error, never a stray
			 * write */
			if (dst < n) {
				kout[dst] = k;
				vout[dst] = s_vals[q];
			} else {
				bad = true;
			}
		}
		__syncthreads();
	}
	if (bad) set_error(ctl, PP_LIBRARY_ERROR, ctl->tick, -1);
}

/* --------------------------------------------------------- bucket sort */

/* Agents move at most one cell per iteration, so the sort after the first
 * iteration needs no general sort. The cells of every replica's grid rows are
 * cut into blocks of W = 2^S cells; bucket (replica, row y, block j) holds the
 * cells [x0, x1) of row y. The moved agents lie in the order of their cells
 * before the move, so the agents that can land in bucket (y, j) are those of
 * three old buckets, (y-1, j), (y, j) and (y+1, j), plus those of the two
 * cells just left and right of the block in row y (rows and columns wrap on
 * the torus): five contiguous ranges of the moved order. The emit counts, per
 * bucket, the moved agents by old cell (ocnt) and the survivors by new cell
 * (ncnt) in three levels of 64-ary sums; the prefix sums of both give the
 * start positions. Each bucket's group sums the levels itself, or, from
 * BPRE_MIN buckets, k_bprefix writes the prefix sums of all buckets and
 * k_brec a record per bucket with its ranges already ordered.
 * k_bucket_sort gives every bucket one work-group, which
 * gathers the agents of its ranges that land in it, in range order (= agent
 * id order), sorts them stably by cell in LDS (a counting sort) and writes
 * them at the bucket's start: the order of the key (cell << 32 | id), as the
 * radix sort gives. */

constexpr uint32_t BCAP = 1024;  /* agents of one bucket sorted in LDS */
constexpr uint32_t NXCD = 8;     /* accelerator dies (XCDs) of the GPU */
constexpr uint32_t W_MAX = 2048; /* cells of a bucket block: 2^S, S <= 11 */
static_assert(W_MAX <= 2 * BCAP, "a dense bucket's counters reuse the staging");
constexpr uint32_t BG_IPT = 6;   /* source positions per thread per round (1536: the three bucket rows of 12800^2) */
static_assert(NWAVE == 4, "the gather reads the waves' counts as one uint4");
/* The automatic choice (sort mode 1): one replica sorts by buckets at any
 * size (12800 x 12800: 89,600 buckets, 0.28 ms against 1.15 ms for four
 * radix passes); a batch of several sorts by buckets while it has at most
 * BUCKET_AUTO_MAX buckets, or at least BUCKET_AUTO_DENSITY agents a bucket
 * on average, else by the radix sort. Each bucket costs a few dependent
 * round trips to memory whatever it holds, so sparse buckets (many small
 * replicas, or sparse populations) favour the radix sort; dense ones do not
 * (30 replicas of 1600 x 1600, 48,000 buckets of about 390 agents: 2.48 s
 * against 3.47 s by radix for 4000 iterations). */
constexpr uint64_t BUCKET_AUTO_MAX = 16384;
constexpr uint64_t BUCKET_AUTO_DENSITY = 64;

/* Clears the count tables, all levels (a tick sorted by the radix sort after
 * a change of the bucket layout: what the previous emit counted is void). */
__global__ __launch_bounds__(WG) void k_bscan_clear(uint32_t *ocnt,
	uint32_t *ncnt, uint32_t words, const Ctl *ctl) {
	if (ctl->error) return;
	for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < words; i += gridDim.x * WG) {
		ocnt[i] = 0;
		ncnt[i] = 0;
	}
}

/* Prefix tables of a large layout (at least BPRE_MIN buckets; with the
 * bucket records below, one replica per iteration: 6400 x 6400, 25,600
 * buckets, 286 -> 260 us; 3200 x 3200, 6400 buckets, 116.6 -> 114.7 us;
 * 2400 x 2400, 4800 buckets, 84.5 -> 85.0 us; 1600 x 1600, 46.5 -> 49.5 us:
 * two more launches outweigh what smaller layouts save), after the old-
 * cell table's words: opre[b], the moved agents of buckets [0, b), and npre[b],
 * the survivors bound for buckets [0, b), for b in [0, NB]. k_bprefix writes
 * them between the emit and the bucket sort, which then reads a range's
 * bounds with two loads instead of summing up to 3 x 63 entries of the count
 * levels in every bucket. */
constexpr uint64_t BPRE_MIN = 6144;
__host__ __device__ __forceinline__ uint32_t opre_off(uint32_t nbc) {
	return cnt_words(nbc);
}
__host__ __device__ __forceinline__ uint32_t npre_off(uint32_t nbc) {
	return cnt_words(nbc) + nbc + 1;
}

/* One wave per table (old cells, new cells) and block of 64 buckets: the sum
 * of the buckets before the block from the count levels 1 and 2 (one load a
 * lane, one round), then the block's own from a wave scan of level 0. */
__global__ __launch_bounds__(WG) void k_bprefix(const uint32_t *ocnt,
	const uint32_t *ncnt, uint32_t *pre, const Ctl *ctl, Params P) {
	const CtlHead H = ctl_head(ctl);
	if (H.error) return;
	const uint32_t NB = H.nbuckets;
	const uint32_t wv = blockIdx.x * NWAVE + threadIdx.x / WAVE;
	const uint32_t q1 = wv >> 1, t = wv & 1u;
	if (q1 > (NB >> 6)) return;
	const uint32_t lane = threadIdx.x & (WAVE - 1);
	const uint32_t *cnt = t ? ncnt : ocnt;
	uint32_t before = 0;
	if (P.cnt_top >= 1) {
		const uint32_t q2 = q1 >> 6;
		if (P.cnt_top >= 2 && lane < q2) before += cnt[cnt_off(P.nb_cap, 2) + lane];
		if (lane < (q1 & 63u)) before += cnt[cnt_off(P.nb_cap, 1) + (q2 << 6) + lane];
	}
	before = wave_sum32(before);
	const uint32_t bk = (q1 << 6) + lane;
	const uint32_t v = bk < NB ? cnt[bk] : 0u;
	const uint32_t excl = before + wave_incl_scan32(v) - v;
	if (bk <= NB) pre[(t ? npre_off(P.nb_cap) : opre_off(P.nb_cap)) + bk] = excl;
}

/* Bucket records of a large layout, after the prefix tables (64-byte
 * aligned): 16 words per bucket that k_brec writes and the bucket sort reads
 * with one load a lane, in place of working out its geometry and ordering its
 * source ranges itself: the cells [cs, ce), the output start and count, the
 * gathered sequence's length and range count, and per range r its first
 * sequence index cum[r] (cum[0] = 0; unused: all ones) and the position
 * minus sequence index delta[r] (unused: 0). */
enum { BR_CS = 0, BR_CE, BR_OUT0, BR_EXPECT, BR_NSRC, BR_NR, BR_CUM1,
	BR_DELTA0 = BR_CUM1 + 4, BR_WORDS = 16 };
static_assert(BR_DELTA0 + 5 <= BR_WORDS, "record layout");
__host__ __device__ __forceinline__ uint32_t brec_off(uint32_t nbc) {
	return (npre_off(nbc) + nbc + 1 + BR_WORDS - 1) & ~(uint32_t)(BR_WORDS - 1);
}

/* One thread per bucket: the bucket's replica, grid rows and block, the
 * buckets its agents come from (the rows above and below, its own, and for
 * rows wider than a block the cells just left and right of the block), their
 * ranges of the moved order from the prefix tables, ordered by start without
 * empty ones or repeats (a sorting network), written as the bucket's record.
 * The layout rules are those of bucket_of and the emit's edge counts. */
__global__ __launch_bounds__(WG) void k_brec(uint32_t *ocnt, const uint32_t *ncnt,
	const Ctl *ctl, Tables T, Params P) {
	const CtlHead H = ctl_head(ctl);
	if (H.error) return;
	const uint32_t NB = H.nbuckets, nrep = H.nrep, S = P.bshift;
	const uint32_t W = 1u << S;
	const uint32_t b = blockIdx.x * WG + threadIdx.x;
	if (b >= NB) return;
	uint32_t slot = P.one_slot, bb = 0;
	if (nrep != 1) {
		const uint32_t lo = ncnt[cnt_words(P.nb_cap) + b];
		slot = T.order[lo];
		bb = T.bbase[lo];
	}
	const Rep &rp = T.rep[slot];
	const uint32_t gx = rp.gx, gy = rp.gy, cb = rp.cb;
	const uint32_t nb = row_blocks(gx, S), h = rows_per_bucket(gx, S);
	const uint32_t nby = (gy + h - 1) / h;
	const uint32_t local = b - bb;
	const uint32_t y = nb == 1 ? local : local / nb, j = local - y * nb;
	const uint32_t x0 = j << S, x1 = min(x0 + W, gx);
	const uint32_t xl = x0 ? x0 - 1 : gx - 1, xr = x1 < gx ? x1 : 0;
	const uint32_t ym = y ? y - 1 : nby - 1, yp = y + 1 < nby ? y + 1 : 0;
	const uint32_t need[5] = {bb + ym * nb + j, b, bb + yp * nb + j,
		bb + y * nb + (xl >> S), bb + y * nb + (xr >> S)};
	const uint32_t row0 = cb + y * h * gx;
	const uint32_t *opre = ocnt + opre_off(P.nb_cap), *npre = ocnt + npre_off(P.nb_cap);
	uint32_t rs[5], re[5];
#pragma unroll
	for (int a = 0; a < 5; a++) {
		rs[a] = opre[need[a]];
		re[a] = opre[need[a] + 1];
	}
	if (nb >= 2) {
		rs[3] = re[3] - ocnt[edge_last_off(P.nb_cap) + need[3]];
		re[4] = rs[4] + ocnt[edge_first_off(P.nb_cap) + need[4]];
	} else {
		rs[3] = re[3] = rs[4] = re[4] = 0xFFFFFFFFu;
	}
	auto ce2 = [&](int a, int c) {
		if (rs[c] < rs[a]) {
			const uint32_t ts = rs[a], te = re[a];
			rs[a] = rs[c]; re[a] = re[c];
			rs[c] = ts; re[c] = te;
		}
	};
	ce2(0, 3); ce2(1, 4); ce2(0, 2); ce2(1, 3); ce2(0, 1);
	ce2(2, 4); ce2(1, 2); ce2(3, 4); ce2(2, 3);
	uint32_t r[BR_WORDS];
#pragma unroll
	for (int i = 0; i < BR_WORDS; i++) r[i] = 0;
#pragma unroll
	for (int i = 0; i < 4; i++) r[BR_CUM1 + i] = 0xFFFFFFFFu;
	uint32_t m = 0, last = 0xFFFFFFFFu, seq = 0;
#pragma unroll
	for (int a = 0; a < 5; a++) {
		if (rs[a] >= re[a] || rs[a] == last) continue;
		/* (m <= a: constant indices after unrolling select, not index) */
#pragma unroll
		for (int c = 0; c <= a; c++) {
			if ((uint32_t)c != m) continue;
			if (c > 0) r[BR_CUM1 + c - 1] = seq;
			r[BR_DELTA0 + c] = rs[a] - seq;
		}
		seq += re[a] - rs[a];
		last = rs[a];
		m++;
	}
	r[BR_CS] = row0 + x0;
	r[BR_CE] = h == 1 ? row0 + x1 : cb + min(y * h + h, gy) * gx;
	r[BR_OUT0] = npre[b];
	r[BR_EXPECT] = npre[b + 1] - npre[b];
	r[BR_NSRC] = seq;
	r[BR_NR] = m;
	uint4 *dst = reinterpret_cast<uint4 *>(ocnt + brec_off(P.nb_cap) + b * BR_WORDS);
#pragma unroll
	for (int i = 0; i < BR_WORDS / 4; i++)
		dst[i] = make_uint4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
}

/* Stable rank of digit d (0 .. nd; nd = no element) among the lanes of the
 * wave: peers with the same digit, and this lane's rank among them. */
__device__ __forceinline__ void wave_match(uint32_t d, uint32_t bits,
	uint32_t &rank, uint32_t &count) {
	const int lane = threadIdx.x & (WAVE - 1);
	uint64_t peers = ~0ull;
	for (uint32_t b = 0; b <= bits; b++) {
		const bool bit = (d >> b) & 1u;
		const uint64_t bal = __ballot(bit);
		peers &= bit ? bal : ~bal;
	}
	rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
	count = (uint32_t)__popcll(peers);
}

/* PRE: the range bounds come from the prefix tables (k_bprefix ran before);
 * else from the count levels. */
template <typename GT, bool PRE>
__global__ __launch_bounds__(WG, 8) void k_bucket_sort(
	const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
	const uint32_t *__restrict__ ocnt, const uint32_t *__restrict__ ncnt,
	uint32_t *__restrict__ okeys, uint32_t *__restrict__ ovals, Ctl *ctl,
	Tables T, Params P, uint32_t *__restrict__ act_cnt, uint32_t act_words,
	GT *grass_at) {
	/* cell offsets | values, in gather order (a dense bucket's per-cell
	 * counters instead) */
	__shared__ uint32_t s_buf[2 * BCAP];
	__shared__ uint16_t s_ix[BCAP];        /* gather index of each sorted slot */
	/* per cell, 16 bits: count, start, end */
	__shared__ __align__(16) uint32_t s_cnt2[W_MAX / 2];
	uint32_t *s_ka = s_buf, *s_va = s_buf + BCAP;
	__shared__ uint32_t s_need[2][BR_WORDS];
	__shared__ uint4 s_wc[2][BG_IPT];     /* per element row: the waves' counts */
	__shared__ uint32_t s_rng[10];
	__shared__ uint32_t s_nrng;
	/* per range: first sequence index (cum) and position - index (delta) */
	__shared__ uint32_t s_cum[5], s_delta[5], s_nsrc;
	__shared__ uint32_t s_tmp[NWAVE];
	const CtlHead H = ctl_head(ctl);
	if (H.error) return;
	const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
	const uint32_t tick = H.tick;
	/* The tick's first kernel: the last move's survivors into live (by the
	 * last work-group, which at small sizes has no bucket), and clear the
	 * action tiles' counters */
	if (blockIdx.x == gridDim.x - 1 && tid < WAVE) fold_live(ctl, P.part, tick);
	for (uint32_t i = blockIdx.x * WG + tid; i < act_words; i += gridDim.x * WG)
		act_cnt[i] = 0;
	const uint32_t NB = H.nbuckets, nrep = H.nrep, S = P.bshift;
	const uint32_t W = 1u << S;

	/* a plain simulation's one replica and its bucket layout, once */
	uint32_t gx1 = 0, gy1 = 0, cb1 = 0, nb1 = 1, h1 = 1, nby1 = 1;
	if (nrep == 1) {
		const Rep &rp = T.rep[P.one_slot];
		gx1 = __builtin_amdgcn_readfirstlane(rp.gx);
		gy1 = __builtin_amdgcn_readfirstlane(rp.gy);
		cb1 = __builtin_amdgcn_readfirstlane(rp.cb);
		nb1 = row_blocks(gx1, S);
		h1 = rows_per_bucket(gx1, S);
		nby1 = (gy1 + h1 - 1) / h1;
	}
	/* Geometry of bucket b: its replica (the largest k with bbase[k] <= b,
	 * from the table after the survivor counts), grid
	 * rows [y h, y h + h), block j (nb >= 2: h = 1), and the buckets its
	 * agents come from: the bucket rows above and below, its own, and
	 * (nb >= 2) those of the cells left and right of the block. */
	struct Geo {
		uint32_t gx, gy, cb, nb, h, y, j;
		uint32_t need[6];
	};
	auto geo = [&](uint32_t b) {
		Geo g;
		g.gx = gx1;
		g.gy = gy1;
		g.cb = cb1;
		g.nb = nb1;
		g.h = h1;
		uint32_t nby = nby1;
		uint32_t bb = 0;
		if (nrep != 1) {
			const uint32_t lo = ncnt[cnt_words(P.nb_cap) + b];
			const Rep &rp = T.rep[T.order[lo]];
			g.gx = rp.gx;
			g.gy = rp.gy;
			g.cb = rp.cb;
			bb = T.bbase[lo];
			g.nb = row_blocks(g.gx, S);
			g.h = rows_per_bucket(g.gx, S);
			nby = (g.gy + g.h - 1) / g.h;
		}
		const uint32_t local = b - bb;
		g.y = g.nb == 1 ? local : local / g.nb;
		g.j = local - g.y * g.nb;
		const uint32_t x0 = g.j << S, x1 = min(x0 + W, g.gx);
		const uint32_t xl = x0 ? x0 - 1 : g.gx - 1;
		const uint32_t xr = x1 < g.gx ? x1 : 0;
		const uint32_t ym = g.y ? g.y - 1 : nby - 1;
		const uint32_t yp = g.y + 1 < nby ? g.y + 1 : 0;
		g.need[0] = b;
		g.need[1] = bb + ym * g.nb + g.j;
		g.need[2] = b;
		g.need[3] = bb + yp * g.nb + g.j;
		g.need[4] = bb + g.y * g.nb + (xl >> S);
		g.need[5] = bb + g.y * g.nb + (xr >> S);
		return g;
	};
	/* [start, end) of the survivors bound for bucket need[0] (new cells)
	 * and of the moved agents of need[1..5] (old cells), from prefix sums of
	 * the emit's counts; wave w takes needs w and w + 4. Issued in one step
	 * (loads) and published to LDS in another, so that a bucket's loads
	 * overlap the previous bucket's sort. Edge counts: agents on the last
	 * cell of need[4] (left of the block) and on the first of need[5]. */
	uint32_t pl[2][CNT_LEVELS], pc[2], pe[2];
	/* With the records, lane i of wave 0 loads word i of the bucket's. */
	uint32_t r_rec = 0;
	auto issue_rec = [&](uint32_t bk) {
		if (w == 0 && lane < (int)BR_WORDS)
			r_rec = at32(ocnt, brec_off(P.nb_cap) + bk * BR_WORDS + (uint32_t)lane);
	};
	auto publish_rec = [&](uint32_t *dst) {
		if (w == 0 && lane < (int)BR_WORDS) dst[lane] = r_rec;
	};
	auto issue = [&](const Geo &g) {
#pragma unroll
		for (int i = 0; i < 2; i++) {
			const int s = w + NWAVE * i;
			pc[i] = pe[i] = 0;
#pragma unroll
			for (int k = 0; k < CNT_LEVELS; k++) pl[i][k] = 0;
			if (s >= 6) continue;
			/* (constant indices: the array stays in registers) */
			const uint32_t q = s == 0 ? g.need[0] : s == 1 ? g.need[1]
				: s == 2 ? g.need[2] : s == 3 ? g.need[3] : s == 4 ? g.need[4]
				: g.need[5];
			const uint32_t *cnt = s == 0 ? ncnt : ocnt;
			pc[i] = cnt[q];
			if (s == 4) pe[i] = ocnt[edge_last_off(P.nb_cap) + q];
			if (s == 5) pe[i] = ocnt[edge_first_off(P.nb_cap) + q];
#pragma unroll
			for (int k = 0; k < CNT_LEVELS; k++) {
				if ((uint32_t)k > P.cnt_top) break;
				const uint32_t sh = 6 * k;
				const uint32_t base = (q >> (sh + 6)) << 6, lim = (q >> sh) & 63u;
				if ((uint32_t)lane < lim) pl[i][k] = cnt[cnt_off(P.nb_cap, k) + base + lane];
			}
		}
	};
	auto publish = [&](uint32_t *dst) {
		if constexpr (PRE) {
			publish_rec(dst);
			return;
		}
#pragma unroll
		for (int i = 0; i < 2; i++) {
			const int s = w + NWAVE * i;
			if (s >= 6) continue;
			uint32_t v = 0;
#pragma unroll
			for (int k = 0; k < CNT_LEVELS; k++) v += pl[i][k];
			v = wave_sum32(v);
			if (lane == 0) {
				dst[2 * s] = v;
				dst[2 * s + 1] = v + pc[i];
				if (s >= 4) dst[8 + s] = pe[i];
			}
		}
	};

	/* Work-groups are dispatched to the XCDs round-robin by index: the
	 * groups of one XCD take a contiguous share of the buckets, so that a
	 * bucket and the bucket rows above and below it, whose agents it
	 * gathers, are read through the same L2. */
	const uint32_t gpx = gridDim.x / NXCD, xcd = blockIdx.x % NXCD;
	const uint32_t share = (NB + NXCD - 1) / NXCD;
	const uint32_t b_end = min(NB, (xcd + 1) * share);
	uint32_t b = blockIdx.x < gpx * NXCD ? xcd * share + blockIdx.x / NXCD : NB;
	if (b >= b_end) b = NB;
	if (b < NB) {
		if constexpr (PRE) issue_rec(b);
		else issue(geo(b));
		publish(s_need[0]);
	}
	__syncthreads();
	for (uint32_t it = 0; b < NB; b += gpx, it++) {
		const uint32_t *nd = s_need[it & 1];
		uint32_t cs, ce, out0, expect, nr, nsrc;
		uint32_t cum1, cum2, cum3, cum4, d0, d1, d2, d3, d4;
		if constexpr (PRE) {
			/* the bucket's record (k_brec) */
			cs = __builtin_amdgcn_readfirstlane(nd[BR_CS]);
			ce = __builtin_amdgcn_readfirstlane(nd[BR_CE]);
			out0 = __builtin_amdgcn_readfirstlane(nd[BR_OUT0]);
			expect = __builtin_amdgcn_readfirstlane(nd[BR_EXPECT]);
			nsrc = __builtin_amdgcn_readfirstlane(nd[BR_NSRC]);
			nr = __builtin_amdgcn_readfirstlane(nd[BR_NR]);
			cum1 = __builtin_amdgcn_readfirstlane(nd[BR_CUM1]);
			cum2 = __builtin_amdgcn_readfirstlane(nd[BR_CUM1 + 1]);
			cum3 = __builtin_amdgcn_readfirstlane(nd[BR_CUM1 + 2]);
			cum4 = __builtin_amdgcn_readfirstlane(nd[BR_CUM1 + 3]);
			d0 = __builtin_amdgcn_readfirstlane(nd[BR_DELTA0]);
			d1 = __builtin_amdgcn_readfirstlane(nd[BR_DELTA0 + 1]);
			d2 = __builtin_amdgcn_readfirstlane(nd[BR_DELTA0 + 2]);
			d3 = __builtin_amdgcn_readfirstlane(nd[BR_DELTA0 + 3]);
			d4 = __builtin_amdgcn_readfirstlane(nd[BR_DELTA0 + 4]);
			/* the ranges' positions, for a dense bucket's path */
			if (tid < 5 && (uint32_t)tid < nr) {
				const uint32_t c0 = tid ? nd[BR_CUM1 + tid - 1] : 0u;
				const uint32_t c1 = (uint32_t)tid + 1 < nr ? nd[BR_CUM1 + tid] : nsrc;
				s_rng[2 * tid] = c0 + nd[BR_DELTA0 + tid];
				s_rng[2 * tid + 1] = c1 + nd[BR_DELTA0 + tid];
			}
			/* thread t clears counter words 4t .. 4t + 3 (W / 2 is a
			 * multiple of 4), as the scan below reads them (ordered
			 * before the counting by the gather's barriers) */
			if (4u * (uint32_t)tid < W / 2)
				*reinterpret_cast<uint4 *>(s_cnt2 + 4u * (uint32_t)tid) =
					make_uint4(0u, 0u, 0u, 0u);
		} else {
			const Geo g = geo(b);
			const uint32_t x0 = g.j << S, x1 = min(x0 + W, g.gx);
			const uint32_t row0 = g.cb + g.y * g.h * g.gx;
			cs = row0 + x0;
			ce = g.h == 1 ? row0 + x1 : g.cb + min(g.y * g.h + g.h, g.gy) * g.gx;
			out0 = nd[0];
			expect = nd[1] - out0;
			/* source ranges of the moved order in position order (a sorting
			 * network on the starts), without empty ones or repeats */
			if (tid == 0) {
				uint32_t rs[5], re[5];
#pragma unroll
				for (int s = 1; s < 6; s++) {
					rs[s - 1] = nd[2 * s];
					re[s - 1] = nd[2 * s + 1];
				}
				if (g.nb >= 2) {
					rs[3] = re[3] - nd[12];
					re[4] = rs[4] + nd[13];
				} else {
					rs[3] = re[3] = rs[4] = re[4] = 0xFFFFFFFFu;
				}
				auto ce2 = [&](int a, int c) {
					if (rs[c] < rs[a]) {
						const uint32_t ts = rs[a], te = re[a];
						rs[a] = rs[c]; re[a] = re[c];
						rs[c] = ts; re[c] = te;
					}
				};
				ce2(0, 3); ce2(1, 4); ce2(0, 2); ce2(1, 3); ce2(0, 1);
				ce2(2, 4); ce2(1, 2); ce2(3, 4); ce2(2, 3);
				uint32_t m = 0, last = 0xFFFFFFFFu, seq = 0;
#pragma unroll
				for (int a = 0; a < 5; a++) {
					if (rs[a] >= re[a] || rs[a] == last) continue;
					s_rng[2 * m] = rs[a];
					s_rng[2 * m + 1] = re[a];
					s_cum[m] = seq;
					s_delta[m] = rs[a] - seq;
					seq += re[a] - rs[a];
					last = rs[a];
					m++;
				}
				for (uint32_t a = m; a < 5; a++) {
					s_cum[a] = 0xFFFFFFFFu;
					s_delta[a] = 0;
				}
				s_nrng = m;
				s_nsrc = seq;
			}
			/* thread t clears counter words 4t .. 4t + 3 (W / 2 is a
			 * multiple of 4), as the scan below reads them */
			if (4u * (uint32_t)tid < W / 2)
				*reinterpret_cast<uint4 *>(s_cnt2 + 4u * (uint32_t)tid) =
					make_uint4(0u, 0u, 0u, 0u);
			__syncthreads();
			nr = s_nrng;
			nsrc = __builtin_amdgcn_readfirstlane(s_nsrc);
			/* range r of sequence index e: the last r with cum[r] <= e (uniform
			 * bounds, so a few compares and selects per element) */
			cum1 = __builtin_amdgcn_readfirstlane(s_cum[1]);
			cum2 = __builtin_amdgcn_readfirstlane(s_cum[2]);
			cum3 = __builtin_amdgcn_readfirstlane(s_cum[3]);
			cum4 = __builtin_amdgcn_readfirstlane(s_cum[4]);
			d0 = __builtin_amdgcn_readfirstlane(s_delta[0]);
			d1 = __builtin_amdgcn_readfirstlane(s_delta[1]);
			d2 = __builtin_amdgcn_readfirstlane(s_delta[2]);
			d3 = __builtin_amdgcn_readfirstlane(s_delta[3]);
			d4 = __builtin_amdgcn_readfirstlane(s_delta[4]);
		}
		/* gather the agents landing in [cs, ce), in range order: the ranges
		 * as one sequence of positions, element e of a round at lane-
		 * consecutive e = v0 + j WG + tid (coalesced), all loads of a round
		 * issued together; the compaction keeps the sequence order */
		uint32_t cnt = 0;
		const uint64_t lt_mask = (1ull << lane) - 1ull;
		for (uint32_t v0 = 0, rnd = 0; v0 < nsrc; v0 += WG * BG_IPT, rnd++) {
			uint32_t kk[BG_IPT], vv[BG_IPT];
#pragma unroll
			for (uint32_t jj = 0; jj < BG_IPT; jj++) {
				const uint32_t e = v0 + jj * WG + tid;
				kk[jj] = DEAD_KEY;
				vv[jj] = 0u;
				if (e < nsrc) {
					const uint32_t d = e >= cum4 ? d4 : e >= cum3 ? d3
						: e >= cum2 ? d2 : e >= cum1 ? d1 : d0;
					const uint32_t p = e + d;
					kk[jj] = at32(keys, p);
					vv[jj] = at32(vals, p);
				}
			}
			/* per element: its rank among the wave's landing ones; per
			 * (element row jj, wave): their count, in LDS */
			uint32_t *wc = reinterpret_cast<uint32_t *>(s_wc[rnd & 1]);
			uint32_t rank[BG_IPT], in_mask = 0;
#pragma unroll
			for (uint32_t jj = 0; jj < BG_IPT; jj++) {
				const bool in = kk[jj] >= cs && kk[jj] < ce;
				const uint64_t bal = __ballot(in);
				rank[jj] = (uint32_t)__popcll(bal & lt_mask);
				in_mask |= in ? 1u << jj : 0u;
				if (lane == 0) wc[jj * NWAVE + w] = (uint32_t)__popcll(bal);
			}
			__syncthreads();
			uint32_t run = 0;
#pragma unroll
			for (uint32_t jj = 0; jj < BG_IPT; jj++) {
				const uint4 c4 = s_wc[rnd & 1][jj];
				const uint32_t off = run + (w > 0 ? c4.x : 0u) + (w > 1 ? c4.y : 0u)
					+ (w > 2 ? c4.z : 0u);
				run += c4.x + c4.y + c4.z + c4.w;
				if (in_mask & (1u << jj)) {
					const uint32_t slot = cnt + off + rank[jj];
					if (slot < BCAP) {
						s_ka[slot] = kk[jj] - cs;
						s_va[slot] = vv[jj];
					}
				}
			}
			cnt += run;
		}
		/* the last chunk's LDS writes land before anyone reads them */
		__syncthreads();
		if (cnt != expect) {
			/* the counts do not describe the agents */
			if (tid == 0) set_error(ctl, PP_LIBRARY_ERROR, tick, -2 - (int32_t)b);
			break;
		}
		/* the next bucket's loads, in flight during this one's sort */
		const uint32_t bn = b + gpx < b_end ? b + gpx : NB;
		if (bn < NB) {
			if constexpr (PRE) issue_rec(bn);
			else issue(geo(bn));
		}
		if (cnt <= BCAP) {
			/* counting sort by cell offset in LDS: counts, exclusive scan,
			 * scatter of the gather indices, then each cell's few slots put
			 * back in gather (= id) order. Cell c's counter is half c & 1
			 * of word c / 2 (values stay below 2^16: cnt <= BCAP). */
			for (uint32_t i = tid; i < cnt; i += WG) {
				const uint32_t c = s_ka[i];
				atomicAdd(&s_cnt2[c >> 1], 1u << ((c & 1u) * 16));
			}
			__syncthreads();
			{
				/* thread t: words 4t .. 4t + 3 (one vector; W / 2 is a
				 * multiple of 4, so a thread has all four or none) */
				static_assert(W_MAX / 2 == 4 * WG, "four counter words a thread");
				const uint32_t w0 = 4u * (uint32_t)tid;
				const bool mine = w0 < W / 2;
				uint4 *p4 = reinterpret_cast<uint4 *>(s_cnt2 + w0);
				const uint4 v4 = mine ? *p4 : make_uint4(0u, 0u, 0u, 0u);
				const uint32_t v[4] = {v4.x, v4.y, v4.z, v4.w};
				uint32_t sum = 0;
#pragma unroll
				for (int i = 0; i < 4; i++) sum += (v[i] & 0xFFFFu) + (v[i] >> 16);
				uint32_t tot;
				uint32_t run = wg_excl_scan<WG>(sum, s_tmp, tot);
				uint32_t o[4];
#pragma unroll
				for (int i = 0; i < 4; i++) {
					const uint32_t lo = v[i] & 0xFFFFu;
					o[i] = run | ((run + lo) << 16);
					run += lo + (v[i] >> 16);
				}
				if (mine) *p4 = make_uint4(o[0], o[1], o[2], o[3]);
			}
			__syncthreads();
			for (uint32_t i = tid; i < cnt; i += WG) {
				const uint32_t c = s_ka[i], sh = (c & 1u) * 16;
				const uint32_t old = atomicAdd(&s_cnt2[c >> 1], 1u << sh);
				s_ix[(old >> sh) & 0xFFFFu] = (uint16_t)i;
			}
			__syncthreads();
			if (bn < NB) publish(s_need[(it + 1) & 1]);
			/* Each counter is now the end of its cell's slots, which hold the
			 * cell's agents in scatter order: an agent's place among them is
			 * the number of its cell's agents gathered before it (few: cells
			 * hold a handful of agents), and it is written there directly.
			 * The sorted pairs, and the emit's moved pairs, are stored
			 * non-temporally: nothing in the writing kernel reads them
			 * again (with the emit's: 1600 x 1600 43.3 -> 42.3 us an
			 * iteration, 12800 x 12800 636 -> 631; the act's write-back
			 * the same way gained nothing). */
			for (uint32_t q = tid; q < cnt; q += WG) {
				const uint32_t i = s_ix[q];
				const uint32_t c = s_ka[i];
				const uint32_t e = (s_cnt2[c >> 1] >> ((c & 1u) * 16)) & 0xFFFFu;
				const uint32_t s0 = c ? (s_cnt2[(c - 1) >> 1] >> (((c - 1) & 1u) * 16)) & 0xFFFFu : 0u;
				uint32_t at = s0;
				for (uint32_t a = s0; a < e; a++) at += s_ix[a] < i;
				__builtin_nontemporal_store(cs + c, &at32(okeys, out0 + at));
				__builtin_nontemporal_store(s_va[i], &at32(ovals, out0 + at));
			}
		} else {
			/* More agents than LDS holds (a dense bucket): count per cell,
			 * then one wave places them in range order straight to HBM. */
			uint32_t *s_run = s_buf;     /* W <= W_MAX <= 2 BCAP counters */
			for (uint32_t i = tid; i < W; i += WG) s_run[i] = 0;
			__syncthreads();
			for (uint32_t r = 0; r < nr; r++) {
				const uint32_t r1 = s_rng[2 * r + 1];
				for (uint32_t p = s_rng[2 * r] + tid; p < r1; p += WG) {
					const uint32_t key = keys[p];
					if (key >= cs && key < ce) atomicAdd(&s_run[key - cs], 1u);
				}
			}
			__syncthreads();
			if (w == 0) {
				const uint32_t per = W / WAVE;   /* W >= 64 */
				uint32_t sum = 0;
				for (uint32_t i = 0; i < per; i++) sum += s_run[lane * per + i];
				uint32_t run = wave_incl_scan32(sum) - sum;
				for (uint32_t i = 0; i < per; i++) {
					const uint32_t v = s_run[lane * per + i];
					s_run[lane * per + i] = run;
					run += v;
				}
				__builtin_amdgcn_wave_barrier();
				for (uint32_t r = 0; r < nr; r++) {
					const uint32_t r1 = s_rng[2 * r + 1];
					for (uint32_t p0 = s_rng[2 * r]; p0 < r1; p0 += WAVE) {
						const uint32_t p = p0 + (uint32_t)lane;
						const uint32_t key = p < r1 ? keys[p] : DEAD_KEY;
						const bool in = key >= cs && key < ce;
						const uint32_t off = in ? key - cs : W;
						uint32_t rank, count;
						wave_match(off, S, rank, count);
						if (in) {
							const uint32_t dst = s_run[off] + rank;
							if (dst < expect) {
								__builtin_nontemporal_store(key, &okeys[out0 + dst]);
								__builtin_nontemporal_store(vals[p], &ovals[out0 + dst]);
							} else {
								set_error(ctl, PP_LIBRARY_ERROR, tick,
									-(int32_t)(1u << 30) - (int32_t)b);
							}
						}
						__builtin_amdgcn_wave_barrier();
						if (in && rank == 0) s_run[off] += count;
						__builtin_amdgcn_wave_barrier();
					}
				}
			}
			if (bn < NB) publish(s_need[(it + 1) & 1]);
		}
		__syncthreads();
		if (bn >= NB) break;
	}
	/* Re-stamp the alive cells of grass slice (t mod REFRESH) with t (keeps
	 * stored regrow ticks within the modular window; the state does not
	 * change). At the end, off the buckets' path; the slice is dealt out
	 * from the groups that had no bucket, if any, and a thread issues its
	 * loads before its stores. */
	{
		const uint32_t per = (P.C + REFRESH - 1) / REFRESH;
		const uint32_t c0 = (tick % REFRESH) * per;
		const uint32_t c1 = min(c0 + per, P.C);
		const uint32_t busy = min(NB, gridDim.x);
		const uint32_t g = (blockIdx.x + gridDim.x - busy) % gridDim.x;
		const uint32_t stride = gridDim.x * WG;
		constexpr int RB = 8;   /* cells a thread loads before it stores */
		for (uint32_t c = c0 + g * WG + tid; c < c1; c += RB * stride) {
			GT gv[RB];
#pragma unroll
			for (int u = 0; u < RB; u++) {
				const uint64_t i = (uint64_t)c + (uint64_t)u * stride;
				gv[u] = i < c1 ? grass_at[i] : (GT)tick;
			}
#pragma unroll
			for (int u = 0; u < RB; u++) {
				const uint64_t i = (uint64_t)c + (uint64_t)u * stride;
				if (i < c1 && grass_alive_at<GT>(gv[u], tick, P.R) && gv[u] != (GT)tick)
					grass_at[i] = (GT)tick;
			}
			if ((uint64_t)c + (uint64_t)RB * stride >= c1) break;
		}
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
		if (q < (uint32_t)A_TILE) return lds_ld(s_val + q);
		return g_val[q];
	}
	__device__ __forceinline__ void set(uint32_t q, uint32_t v) const {
		if (q < (uint32_t)A_TILE) lds_st(s_val + q, v);
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

/* What the actions need of a replica. */
struct RA {
	uint32_t cb, step, R;
	uint32_t sheep_gain, wolves_gain, sheep_thr, wolves_thr;
	uint32_t key0, key1;
};

/* For a replica index that may differ between lanes. */
__device__ __forceinline__ RA load_ra_lane(const Tables &T, uint32_t k,
	uint32_t tick) {
	const Rep &rp = T.rep[T.order[k]];
	RA a;
	a.cb = rp.cb;
	a.step = tick - rp.t0;
	a.R = rp.R;
	a.sheep_gain = rp.sheep_gain;
	a.wolves_gain = rp.wolves_gain;
	a.sheep_thr = rp.sheep_thr;
	a.wolves_thr = rp.wolves_thr;
	a.key0 = rp.key0;
	a.key1 = rp.key1;
	return a;
}

/* For a replica index uniform over the wave (slot_hint: its slot, when
 * known without the table). */
__device__ __forceinline__ RA load_ra(const Tables &T, uint32_t k,
	uint32_t tick, uint32_t slot_hint = NO_REP) {
	/* k is uniform: the fields are moved to scalar registers, so later
	 * uses do not wait on the vector loads (which retire in order with
	 * every other global access). */
	const Rep &rp = T.rep[slot_hint != NO_REP ? slot_hint
		: __builtin_amdgcn_readfirstlane(T.order[k])];
	RA a;
	a.cb = __builtin_amdgcn_readfirstlane(rp.cb);
	a.step = tick - __builtin_amdgcn_readfirstlane(rp.t0);
	a.R = __builtin_amdgcn_readfirstlane(rp.R);
	a.sheep_gain = __builtin_amdgcn_readfirstlane(rp.sheep_gain);
	a.wolves_gain = __builtin_amdgcn_readfirstlane(rp.wolves_gain);
	a.sheep_thr = __builtin_amdgcn_readfirstlane(rp.sheep_thr);
	a.wolves_thr = __builtin_amdgcn_readfirstlane(rp.wolves_thr);
	a.key0 = __builtin_amdgcn_readfirstlane(rp.key0);
	a.key1 = __builtin_amdgcn_readfirstlane(rp.key1);
	return a;
}

/* Adds v to counter idx of the replica with table index k, in an action tile
 * whose positions belong to several replicas (the first with index k_first):
 * LDS slots for the first MIX_SLOTS of them, global atomics beyond. */
__device__ __forceinline__ void mix_add(unsigned long long (*s_mix)[NACC],
	const Tables &T, uint32_t k, uint32_t k_first, int idx,
	unsigned long long v) {
	if (!v) return;
	const uint32_t d = k - k_first;
	if (d < (uint32_t)MIX_SLOTS) lds_add(&s_mix[d][idx], v);
	else glb_add(&T.rctl[T.order[k]].acc[idx], v);
}

/* Adds the per-thread counters of one wave to dst (wave sums, one atomic
 * per counter from lane 0). Kept out of line: it runs when an action
 * work-group moves on to another replica's tiles, and inlined it would hold
 * registers over the whole tile loop. */
__device__ __noinline__ void wave_flush(unsigned long long a0,
	unsigned long long a1, unsigned long long a2, unsigned long long a3,
	unsigned long long a4, unsigned long long a5, unsigned long long a6,
	unsigned long long *dst) {
	unsigned long long v[NACC] = {a0, a1, a2, a3, a4, a5, a6};
	const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
	for (int k = 0; k < NACC; k++) {
		const unsigned long long t = wave_sum(v[k]);
		if (lane == 0 && t) glb_add(&dst[k], t);
	}
}

/* Adds v outputs to action tile t: its count, and the sum of its block of
 * 64 tiles (at l1o), which the scan works from. */
__device__ __forceinline__ void tile_add(uint32_t *act_cnt, uint32_t l1o,
	uint32_t t, uint32_t v) {
	atomicAdd(&act_cnt[t], v);
	atomicAdd(&act_cnt[l1o + (t >> 6)], v);
}

/* Work-group sums of the action counters, added to dst (a replica's acc, or
 * a partial-sum slot):
 * AC_OUT is the sheep plus the wolves emitted; counts are summed in 32 bits
 * and each energy sum as two 32-bit pieces split at bit 24 (exact while
 * every lane's sum is below 2^48; a wave holding a larger one sums it in 64
 * bits). acc is cleared. */
template <int NW>
__device__ __forceinline__ void act_wg_add(unsigned long long (&acc)[NACC],
	unsigned long long (*s_red)[NACC], unsigned long long *dst) {
	const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
	const uint32_t ns = wave_sum32((uint32_t)acc[AC_SHEEP]);
	const uint32_t nw = wave_sum32((uint32_t)acc[AC_WOLVES]);
	const uint32_t ne = wave_sum32((uint32_t)acc[AC_EATEN]);
	const uint32_t no = wave_sum32(lo32(acc[AC_OCC]));
	const uint32_t nv = wave_sum32(hi32(acc[AC_OCC]));
	const uint32_t nb = wave_sum32(hi32(acc[AC_OUT]));
	unsigned long long es, ew;
	if (__ballot(((acc[AC_SHEEP_EN] | acc[AC_WOLVES_EN]) >> 48) != 0) == 0) {
		es = wave_sum32((uint32_t)(acc[AC_SHEEP_EN] & 0xFFFFFFu))
			+ ((unsigned long long)wave_sum32((uint32_t)(acc[AC_SHEEP_EN] >> 24)) << 24);
		ew = wave_sum32((uint32_t)(acc[AC_WOLVES_EN] & 0xFFFFFFu))
			+ ((unsigned long long)wave_sum32((uint32_t)(acc[AC_WOLVES_EN] >> 24)) << 24);
	} else {
		es = wave_sum(acc[AC_SHEEP_EN]);
		ew = wave_sum(acc[AC_WOLVES_EN]);
	}
	if (lane == 0) {
		s_red[w][AC_SHEEP] = ns;
		s_red[w][AC_WOLVES] = nw;
		s_red[w][AC_SHEEP_EN] = es;
		s_red[w][AC_WOLVES_EN] = ew;
		s_red[w][AC_EATEN] = ne;
		s_red[w][AC_OCC] = no + nv * HI32;
		s_red[w][AC_OUT] = (unsigned long long)(ns + nw) + nb * HI32;
	}
#pragma unroll
	for (int k = 0; k < NACC; k++) acc[k] = 0;
	__syncthreads();
	if (threadIdx.x < NACC) {
		unsigned long long t = 0;
		for (int i = 0; i < NW; i++) t += s_red[i][threadIdx.x];
		if (t) atomicAdd(&dst[threadIdx.x], t);
	}
	__syncthreads();
}

/* LDS of one action work-group. */
template <typename GT>
struct ActLds {
	uint32_t key[A_TILE + 1];
	uint32_t val[A_TILE];
	uint32_t slow_lp[A_TILE];
	GT slow_g[A_TILE];
	uint32_t nslow;
	uint32_t tmp[WIDE_WG / WAVE];
	unsigned long long red[WIDE_WG / WAVE][NACC];
	unsigned long long mix[MIX_SLOTS][NACC];
};

/* Agent actions of one tile of A_TILE sorted positions (loaded in S.key /
 * S.val, cnt valid), whose agents belong to the replicas k_first..k_last of
 * the table. MIXED = false: one replica (ra), counters summed into acc;
 * MIXED = true: several, each cell looks up its replica and counters go
 * through LDS slots. See k_act for the rules. Agents whose energy after
 * their action exceeds V_ENERGY are counted in the high half of AC_OCC. */
template <typename GT, bool MIXED, int NT = WG>
__device__ __forceinline__ void act_tile(ActLds<GT> &S, uint32_t tile,
	uint32_t cnt, const uint32_t *__restrict__ keys, uint32_t *vals,
	GT *grass_at, uint32_t *act_cnt, uint32_t l1o, const Tables &T, uint32_t nrep,
	uint32_t tick, uint32_t L, uint32_t Rmax, uint32_t k_first,
	uint32_t k_last, const RA &ra_u, unsigned long long (&acc)[NACC]) {
	const int tid = threadIdx.x;
	constexpr int IPT = A_TILE / NT;
	/* wide groups hold two positions a thread, which leaves the registers
	 * for cells of up to eight in the second path (1600 x 1600: 41.6 ->
	 * 41.4 us an iteration) */
	constexpr uint32_t KR2 = NT > WG ? 8u : K_REG2;
	const uint32_t base = tile * A_TILE;
	const uint32_t alive_at = tick + 1;
	const SegVals sv{S.val, vals + base};
	/* Heads of this thread's positions; their grass loads are issued
	 * together before any segment is processed. */
	uint32_t cj[IPT];
	GT gj[IPT];
	bool hj[IPT];
	/* every lane loads (the tile's first cell where it holds no head), so
	 * the loads are issued unconditionally */
	const uint32_t c_safe = __builtin_amdgcn_readfirstlane(S.key[1]);
#pragma unroll
	for (int j = 0; j < IPT; j++) {
		const uint32_t lp = tid * IPT + j;
		cj[j] = S.key[1 + lp];
		hj[j] = lp < cnt && S.key[lp] != cj[j];
		gj[j] = grass_cell(grass_at, hj[j] ? cj[j] : c_safe);
	}
	/* Cells with one agent (most of them) are handled in place; heads of
	 * longer segments are listed and then processed by consecutive lanes, so
	 * the long serial path runs once per wave, not once per position of every
	 * lane. Segments are independent: the order of the list does not change
	 * any result. In a tile of several replicas every cell takes the general
	 * path. */
	if (tid == 0) S.nslow = 0;
	if (MIXED)
		for (int i = tid; i < MIX_SLOTS * NACC; i += NT) (&S.mix[0][0])[i] = 0;
	__syncthreads();
	/* the grass values (and the next tile's pairs, issued before them) */
	wait_global();
	uint32_t nout = 0;
	/* sums over this thread's single-agent cells: at most IPT outputs
	 * pairs below 2^25 each, so 32 bits hold them */
	uint32_t f_sheep = 0, f_wolves = 0, f_se = 0, f_we = 0, f_eaten = 0,
		f_occ = 0, f_ovf = 0, f_birth = 0;
#pragma unroll
	for (int j = 0; j < IPT; j++) {
		if (!hj[j]) continue;
		const uint32_t lp = tid * IPT + j;
		const uint32_t c = cj[j];
		if (!MIXED && lp + 1 < cnt && S.key[2 + lp] != c) {
			/* A cell with one agent: nothing to shuffle, a lone wolf finds no
			 * sheep, a lone agent is never eaten. */
			const uint32_t v = S.val[lp];
			int32_t e = (int32_t)(v & V_ENERGY);
			const uint32_t t = (v & V_WOLF) ? 1u : 0u;
			uint32_t thr = ra_u.wolves_thr;
			uint32_t ate = 0;
			if (!t) {
				thr = ra_u.sheep_thr;
				if (grass_alive_at<GT>(gj[j], alive_at, Rmax)) {
					e += (int32_t)ra_u.sheep_gain;
					grass_cell(grass_at, c) = (GT)(tick + 1 + ra_u.R);
					ate = 1;
				}
			}
			uint32_t nv = v & V_WOLF;
			uint32_t ce = 0;
			if (e > (int32_t)thr && (v & V_REPRO)) {
				nv |= V_CHILD | ((e & 1) ? V_ODD : 0u);
				ce = (uint32_t)(e / 2);
				e -= e / 2;
			}
			f_ovf += e > (int32_t)V_ENERGY;
			S.val[lp] = nv | ((uint32_t)e & V_ENERGY);
			const uint32_t outs = (nv & V_CHILD) ? 2u : 1u;
			const uint32_t en = (uint32_t)e + ce;
			if (t) {
				f_wolves += outs;
				f_we += en;
			} else {
				f_sheep += outs;
				f_se += en;
			}
			f_eaten += ate;
			f_occ += 1;
			f_birth += outs - 1;
			nout += outs;
			continue;
		}
		const uint32_t slot = atomicAdd(&S.nslow, 1u);
		S.slow_lp[slot] = lp;
		S.slow_g[slot] = gj[j];
	}
	if (!MIXED) {
		acc[AC_SHEEP] += f_sheep;
		acc[AC_WOLVES] += f_wolves;
		acc[AC_SHEEP_EN] += f_se;
		acc[AC_WOLVES_EN] += f_we;
		acc[AC_EATEN] += f_eaten;
		acc[AC_OCC] += f_occ + f_ovf * HI32;
		acc[AC_OUT] += (f_sheep + f_wolves) + f_birth * HI32;
	}
	__syncthreads();
	const uint32_t nslow = S.nslow;
#pragma unroll 1
	for (uint32_t si = tid; si < nslow; si += NT) {
		const uint32_t lp = S.slow_lp[si];
		const uint32_t c = S.key[1 + lp];
		const GT gslow = S.slow_g[si];
		const uint32_t kk = MIXED ? rep_index(T, nrep, c) : k_first;
		const RA ra = MIXED ? load_ra_lane(T, kk, tick) : ra_u;
		const uint32_t end = seg_end(S.key, keys, lp, cnt, base, L, c);
		const uint32_t k = end - lp;
		unsigned long long sc[NACC] = {0, 0, 0, 0, 0, 0, 0};
		sc[AC_OCC] = 1;
		/* A cell of at most KR agents in registers, with the same steps as
		 * the general path below: Philox block 0 covers the k - 1 <= 4
		 * draws of a cell of up to five, block 1 the rest (KR up to 8).
		 * The Fisher-Yates swaps act on a permutation of 4-bit indices, and
		 * the values are read from the tile already in acting order; sheep
		 * and eaten agents are bit masks over acting positions. */
		auto reg_cell = [&](auto kr) {
			constexpr uint32_t KR = decltype(kr)::value;
			static_assert(KR <= 8, "4-bit indices in one word");
			uint32_t perm = 0x76543210u;
			/* the swaps of draws m0..m0+3 (i descending across both calls;
			 * one block live at a time) */
			auto swaps = [&](const uint4 w, uint32_t m0) {
#pragma unroll
				for (int i = KR - 1; i >= 1; i--) {
					if ((uint32_t)i >= k) continue;
					const uint32_t m = k - 1 - (uint32_t)i - m0;
					if constexpr (KR > 5)
						if (m >= 4) continue;   /* another block's (m0 <= m' < m0+4) */
					const uint32_t word = m == 0 ? w.x : m == 1 ? w.y : m == 2 ? w.z : w.w;
					const uint32_t jj = pp_bounded(word, (uint32_t)i + 1);
					const uint32_t x = ((perm >> (4 * i)) ^ (perm >> (4 * jj))) & 0xFu;
					perm ^= (x << (4 * i)) | (x << (4 * jj));
				}
			};
			swaps(pp_philox(c - ra.cb, ra.step, PP_STREAM_SHUFFLE, 0, ra.key0, ra.key1), 0);
			if constexpr (KR > 5)
				swaps(pp_philox(c - ra.cb, ra.step, PP_STREAM_SHUFFLE, 1, ra.key0, ra.key1),
					4);
			uint32_t V[KR];
			uint32_t sheep = 0;
#pragma unroll
			for (uint32_t q = 0; q < KR; q++) {
				V[q] = q < k ? sv.get(lp + ((perm >> (4 * q)) & 0xFu)) : 0u;
				if (q < k && !(V[q] & V_WOLF)) sheep |= 1u << q;
			}
			bool grass = grass_alive_at<GT>(gslow, alive_at, Rmax);
			bool ate = false;
			uint32_t killed = 0, ovf = 0;
#pragma unroll
			for (uint32_t q = 0; q < KR; q++) {
				if (q >= k || ((killed >> q) & 1u)) continue;
				const uint32_t v = V[q];
				int32_t e = (int32_t)(v & V_ENERGY);
				uint32_t thr;
				if (!(v & V_WOLF)) {
					if (grass) {
						grass = false;
						ate = true;
						e += (int32_t)ra.sheep_gain;
					}
					thr = ra.sheep_thr;
				} else {
					/* the first sheep in acting order not yet eaten */
					const uint32_t prey = sheep & ~killed;
					if (prey) {
						killed |= prey & (0u - prey);
						e += (int32_t)ra.wolves_gain;
					}
					thr = ra.wolves_thr;
				}
				uint32_t nv = v & V_WOLF;
				if (e > (int32_t)thr && (v & V_REPRO)) {
					nv |= V_CHILD | ((e & 1) ? V_ODD : 0u);
					e -= e / 2;
				}
				ovf += e > (int32_t)V_ENERGY;
				V[q] = nv | ((uint32_t)e & V_ENERGY);
			}
			if (ate) grass_cell(grass_at, c) = (GT)(tick + 1 + ra.R);
			/* per-cell sums fit 32 bits: at most 2 KR outputs below
			 * 2^24 each */
			uint32_t n_sheep = 0, n_wolves = 0, e_sheep = 0, e_wolves = 0,
				births = 0;
#pragma unroll
			for (uint32_t q = 0; q < KR; q++) {
				if (q >= k) continue;
				const uint32_t v = V[q] | (((killed >> q) & 1u) ? V_KILLED : 0u);
				sv.set(lp + q, v);
				const uint32_t e = v & V_ENERGY;
				const uint32_t outs = ((v & V_KILLED) == 0) + ((v & V_CHILD) != 0);
				const uint32_t ea = ((v & V_KILLED) ? 0u : e)
					+ ((v & V_CHILD) ? ((v & V_ODD) ? e - 1 : e) : 0u);
				if (v & V_WOLF) {
					n_wolves += outs;
					e_wolves += ea;
				} else {
					n_sheep += outs;
					e_sheep += ea;
				}
				births += (v & V_CHILD) != 0;
				if (lp + q < (uint32_t)A_TILE) nout += outs;
				else if (outs) tile_add(act_cnt, l1o, tile + (lp + q) / A_TILE, outs);
			}
			sc[AC_SHEEP] = n_sheep;
			sc[AC_WOLVES] = n_wolves;
			sc[AC_SHEEP_EN] = e_sheep;
			sc[AC_WOLVES_EN] = e_wolves;
			sc[AC_OUT] = (n_sheep + n_wolves) + births * HI32;
			sc[AC_EATEN] = ate ? 1 : 0;
			sc[AC_OCC] = 1 + ovf * HI32;
		};
		if (k <= K_REG) {
			reg_cell(KConst<K_REG>{});
		} else if (k <= KR2) {
			/* (a wave takes this branch only for its rare cells of six, or
			 * six to eight in wide groups) */
			reg_cell(KConst<KR2>{});
		} else {
			if (k > 1) {
				uint4 r = make_uint4(0, 0, 0, 0);
				for (uint32_t i = k - 1, m = 0; i > 0; i--, m++) {
					if ((m & 3u) == 0)
						r = pp_philox(c - ra.cb, ra.step, PP_STREAM_SHUFFLE, m >> 2,
							ra.key0, ra.key1);
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
			bool grass = grass_alive_at<GT>(gslow, alive_at, Rmax);
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
						e += (int32_t)ra.sheep_gain;
					}
					thr = ra.sheep_thr;
				} else {
					for (uint32_t o = lp; o < end; o++) {
						const uint32_t u = sv.get(o);
						if (!(u & (V_WOLF | V_KILLED))) {
							sv.set(o, u | V_KILLED);
							e += (int32_t)ra.wolves_gain;
							break;
						}
					}
					thr = ra.wolves_thr;
				}
				uint32_t nv = v & V_WOLF;
				if (e > (int32_t)thr && (v & V_REPRO)) {
					nv |= V_CHILD | ((e & 1) ? V_ODD : 0u);
					e -= e / 2;
				}
				if (e > (int32_t)V_ENERGY) sc[AC_OCC] += HI32;
				sv.set(q, nv | ((uint32_t)e & V_ENERGY));
			}
			if (ate) grass_cell(grass_at, c) = (GT)(tick + 1 + ra.R);
			sc[AC_EATEN] = ate ? 1 : 0;
			for (uint32_t q = lp; q < end; q++) {
				const uint32_t v = sv.get(q);
				const uint32_t e = v & V_ENERGY;
				const uint32_t t = (v & V_WOLF) ? 1u : 0u;
				const uint32_t outs = ((v & V_KILLED) == 0) + ((v & V_CHILD) != 0);
				const uint32_t ea = ((v & V_KILLED) ? 0u : e)
					+ ((v & V_CHILD) ? ((v & V_ODD) ? e - 1 : e) : 0u);
				if (t) {
					sc[AC_WOLVES] += outs;
					sc[AC_WOLVES_EN] += ea;
				} else {
					sc[AC_SHEEP] += outs;
					sc[AC_SHEEP_EN] += ea;
				}
				sc[AC_OUT] += outs + ((v & V_CHILD) ? HI32 : 0ull);
				if (q < (uint32_t)A_TILE) nout += outs;
				else if (outs) tile_add(act_cnt, l1o, tile + q / A_TILE, outs);
			}
		}
		if (!MIXED) {
#pragma unroll
			for (int i = 0; i < NACC; i++) acc[i] += sc[i];
		} else {
#pragma unroll
			for (int i = 0; i < NACC; i++) mix_add(S.mix, T, kk, k_first, i, sc[i]);
		}
	}
	__syncthreads();
	/* Write back the positions this tile owns (not those of the previous
	 * tile's last segment, which its lane wrote). */
	for (uint32_t i = tid; i < cnt; i += NT)
		if (S.key[1 + i] != S.key[0]) at32(vals, base + i) = S.val[i];
	uint32_t total;
	(void)wg_excl_scan<NT>(nout, S.tmp, total);
	if (tid == 0 && total) tile_add(act_cnt, l1o, tile, total);
	if (MIXED) {
		const uint32_t nk = min(k_last - k_first + 1, (uint32_t)MIX_SLOTS);
		for (uint32_t i = tid; i < nk * NACC; i += NT) {
			const unsigned long long v = (&S.mix[0][0])[i];
			if (v)
				atomicAdd(&T.rctl[T.order[k_first + i / NACC]].acc[i % NACC], v);
		}
		__syncthreads();
	}
}

/* Agent actions for the agents sorted by cell (ascending agent id within a
 * cell). A work-group stages A_TILE sorted positions in LDS; the thread owning
 * a position that starts a run of equal cells (the segment that find_cell_idx,
 * pp_gpu.cl:834-867, delimits) acts for the whole cell in sequence:
 *   1. Fisher-Yates permutation of the segment, the draw for position i being
 *      word m%4 of Philox block (local cell, local iteration, SHUFFLE, m/4),
 *      m = k-1-i (random acting order: Cell.java:246, pp_cpu.cl:236-259);
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
 * newborn; act_cnt[t] (zeroed by the tick's first kernel) receives the
 * number of agents the positions of tile t emit, and the sum of its block of
 * 64 tiles the same. Per replica, the counts and energy
 * sums of the new state, the grass eaten, the occupied cells and the outputs
 * are summed into its control words, in registers over the tiles of one
 * replica. Tiles holding agents of several replicas are listed for
 * k_act_mixed (segments never span replicas, so the two kernels touch
 * disjoint cells and values). */
template <typename GT, int NT>
__global__ __launch_bounds__(NT, 6) void k_act(const uint32_t *__restrict__ keys,
	uint32_t *vals, GT *grass_at, uint32_t *act_cnt, uint32_t *mixed_list,
	uint32_t *ocnt, uint32_t *ncnt, Ctl *ctl, Tables T, Params P) {
	constexpr int IPT = A_TILE / NT;
	__shared__ ActLds<GT> S;
	const CtlHead H = ctl_head(ctl);
	if (H.error) return;
	const int tid = threadIdx.x;
	const uint32_t tick = H.tick, nrep = H.nrep;
	/* The sort has read the bucket counts: clear the entries in use for
	 * this tick's emit. */
	if (P.bucket_on) {
		const uint32_t NB = H.nbuckets;
		for (uint32_t k = 0; k <= P.cnt_top; k++) {
			const uint32_t o = cnt_off(P.nb_cap, (int)k), n = (NB >> (6 * k)) + 1;
			for (uint32_t i = blockIdx.x * NT + tid; i < n; i += gridDim.x * NT) {
				ocnt[o + i] = 0;
				ncnt[o + i] = 0;
			}
		}
		for (uint32_t i = blockIdx.x * NT + tid; i <= NB; i += gridDim.x * NT) {
			ocnt[edge_first_off(P.nb_cap) + i] = 0;
			ocnt[edge_last_off(P.nb_cap) + i] = 0;
		}
	}
	/* one replica: the emit ends the tick (no k_act_scan) and moves for
	 * the next; it reads that tick here */
	if (P.solo && blockIdx.x == 0 && tid == 0) ctl->emit_tick = tick + 1;
	const uint32_t L = H.live(tick);
	const uint32_t ntiles = (L + A_TILE - 1) / A_TILE;
	unsigned long long acc[NACC] = {0, 0, 0, 0, 0, 0, 0};
	uint32_t acc_k = NO_REP;   /* replica the registers sum for */
	RA ra{};
	/* Each group takes a contiguous run of tiles, so it crosses few replica
	 * boundaries. The next tile's pairs are loaded into registers before
	 * this tile's stores are issued: the loads then do not wait behind
	 * them. (The values of a segment running on from this tile into the
	 * next are read stale there, but never used: that tile neither acts for
	 * nor writes back the positions of an earlier tile's segment.) */
	const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
	const uint32_t t_beg = blockIdx.x * per;
	const uint32_t t_end = min(ntiles, t_beg + per);
	uint32_t pk[IPT], pv[IPT], pprev = DEAD_KEY;
	auto fetch = [&](uint32_t t) {
		const uint32_t base = t * A_TILE;
#pragma unroll
		for (int j = 0; j < IPT; j++) {
			const uint32_t i = base + tid + j * NT;
			pk[j] = DEAD_KEY;
			pv[j] = 0u;
			if (i < L) {
				pk[j] = at32(keys, i);
				pv[j] = at32(vals, i);
			}
		}
		if (tid == 0) pprev = base ? keys[base - 1] : DEAD_KEY;
	};
	if (t_beg < t_end) {
		fetch(t_beg);
		/* A batch of one replica (every plain simulation): its constants
		 * before the first tile. */
		if (nrep == 1) {
			acc_k = 0;
			ra = load_ra(T, 0, tick, P.one_slot);
		}
	}
	wait_global();

	for (uint32_t tile = t_beg; tile < t_end; tile++) {
		const uint32_t cnt = min((uint32_t)A_TILE, L - tile * A_TILE);
#pragma unroll
		for (int j = 0; j < IPT; j++) {
			S.key[1 + tid + j * NT] = pk[j];
			S.val[tid + j * NT] = pv[j];
		}
		if (tid == 0) S.key[0] = pprev;
		__syncthreads();
		/* Replicas of this tile: keys ascend, so they are the table entries
		 * k_first..k_last. Every branch below is uniform over the group. */
		const uint32_t k_first = rep_index(T, nrep,
			__builtin_amdgcn_readfirstlane(S.key[1]));
		const uint32_t k_last = rep_index(T, nrep,
			__builtin_amdgcn_readfirstlane(S.key[cnt]));
		/* The next tile's pairs are requested inside each branch, after
		 * every load this tile waits on before its grass, and both branches
		 * end in an explicit wait (act_tile waits on the grass loads issued
		 * after them): the loop's top then never waits for them behind this
		 * tile's stores. */
		if (k_first != k_last) {
			if (tile + 1 < t_end) fetch(tile + 1);
			if (tid == 0) mixed_list[atomicAdd(&ctl->nmixed, 1u)] = tile;
			__syncthreads();
			wait_global();
		} else {
			if (k_first != acc_k) {
				if (acc_k != NO_REP) {
					wave_flush(acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], acc[6],
						&T.rctl[T.order[acc_k]].acc[0]);
#pragma unroll
					for (int i = 0; i < NACC; i++) acc[i] = 0;
				}
				acc_k = k_first;
				ra = load_ra(T, k_first, tick);
			}
			if (tile + 1 < t_end) fetch(tile + 1);
			act_tile<GT, false, NT>(S, tile, cnt, keys, vals, grass_at,
				act_cnt, P.act_l1, T, nrep, tick, L, P.R, k_first, k_last, ra, acc);
		}
	}
	/* one replica: a partial-sum slot, words of the tick's parity (the
	 * emit adds them up) */
	if (acc_k != NO_REP)
		act_wg_add<NT / WAVE>(acc, S.red, nrep == 1
			? P.part + (blockIdx.x % PART_SLOTS) * PART_STRIDE + part_acc(tick)
			: &T.rctl[T.order[acc_k]].acc[0]);
}

/* The tiles k_act listed as holding agents of several replicas (at most one
 * per replica boundary): the same actions with a table lookup per cell. */
template <typename GT>
__global__ __launch_bounds__(WG) void k_act_mixed(
	const uint32_t *__restrict__ keys, uint32_t *vals, GT *grass_at,
	uint32_t *act_cnt, const uint32_t *mixed_list, Ctl *ctl, Tables T,
	Params P) {
	__shared__ ActLds<GT> S;
	if (ctl->error) return;
	const uint32_t tick = ctl->tick, nrep = ctl->nrep;
	const uint32_t L = ctl->live[tick & 1];
	const uint32_t nm = ctl->nmixed;
	unsigned long long acc[NACC] = {0, 0, 0, 0, 0, 0, 0};
	const RA none{};
	for (uint32_t i = blockIdx.x; i < nm; i += gridDim.x) {
		const uint32_t tile = mixed_list[i];
		const uint32_t cnt = load_tile(keys, vals, S.key, S.val, tile * A_TILE, L);
		const uint32_t k_first = rep_index(T, nrep,
			__builtin_amdgcn_readfirstlane(S.key[1]));
		const uint32_t k_last = rep_index(T, nrep,
			__builtin_amdgcn_readfirstlane(S.key[cnt]));
		act_tile<GT, true>(S, tile, cnt, keys, vals, grass_at,
			act_cnt, P.act_l1, T, nrep, tick, L, P.R, k_first, k_last, none, acc);
	}
}

/* A batch of several replicas (one replica: end_tick_solo, in the emit).
 * One work-group: exclusive scan of the output counts of the action tiles'
 * blocks of 64 (the block output offsets; a tile's offset adds the counts of
 * the tiles before it in its block), then the end of tick t for every
 * replica of the batch (in table order, so that its agents' batch positions
 * are the prefix of the outputs of the replicas before it). Grass statistics come from the
 * regrow-tick encoding without touching the grid: with D dead cells and
 * countdown sum S, every dead countdown drops by one (S -= D), the cells whose
 * regrow tick is t+1 come alive (ring slot (t+1) mod (R+1)), and the E cells
 * eaten at this tick restart at R (slot t mod (R+1)): exactly the grass
 * kernel (pp_gpu.cl:448-476) plus the sums of reduce_grass1/2
 * (pp_gpu.cl:485-583). Then the replica's statistics row s+1 is written, its
 * energy limit and capacity checked (pp_gpu.c:847-852) and its counters
 * cleared; a replica that reached its iteration count is done, one with an
 * agent above the energy limit or above its capacity failed (the energy
 * limit first, as the oracle checks), and the emit drops the agents of
 * both. The batch fails when its outputs exceed the agent buffers. */

template <int NT>
__global__ __launch_bounds__(NT) void k_act_scan(
	uint32_t *__restrict__ act_cnt, Ctl *ctl, Tables T, Params P) {
	__shared__ uint32_t s_tmp[NT / WAVE];
	/* the one-wave form: the first 64 block sums (of those allocated),
	 * requested with the control words */
	uint32_t *blk = act_cnt + P.act_l1;
	uint32_t v_first = 0;
	if (NT == WAVE && threadIdx.x < (P.act_l1 + 63) / 64 + 1)
		v_first = blk[threadIdx.x];
	const CtlHead H = ctl_head(ctl);
	if (H.error) return;
	const uint32_t tick = H.tick;
	const uint32_t L = H.live(tick);
	const uint32_t nt = (L + A_TILE - 1) / A_TILE;
	/* the sums of blocks of 64 tiles, scanned in place: the emit adds the
	 * counts of the tiles before its own within the block */
	const uint32_t nbk = (nt + 63) / 64;
	uint32_t total;
	if (NT == WAVE) {
		/* few blocks: one wave, 64 at a time, no barriers */
		const uint32_t lane = threadIdx.x;
		uint32_t carry = 0;
		for (uint32_t c0 = 0; c0 < nbk; c0 += WAVE) {
			const uint32_t q = c0 + lane;
			const uint32_t v = q < nbk ? (c0 == 0 ? v_first : blk[q]) : 0u;
			const uint32_t x = wave_incl_scan32(v);
			if (q < nbk) blk[q] = carry + x - v;
			carry += (uint32_t)__builtin_amdgcn_readlane((int)x, WAVE - 1);
		}
		total = carry;
	} else {
		const uint32_t per = (nbk + NT - 1) / NT;
		const uint32_t t0 = threadIdx.x * per, t1 = min(t0 + per, nbk);
		uint32_t sum = 0;
		for (uint32_t t = t0; t < t1; t += SCAN_UNROLL) {
			uint32_t v[SCAN_UNROLL];
#pragma unroll
			for (int u = 0; u < SCAN_UNROLL; u++)
				v[u] = t + u < t1 ? blk[t + u] : 0u;
#pragma unroll
			for (int u = 0; u < SCAN_UNROLL; u++) sum += v[u];
		}
		uint32_t run = wg_excl_scan<NT>(sum, s_tmp, total);
		if (total <= P.cap) {
			for (uint32_t t = t0; t < t1; t += SCAN_UNROLL) {
				uint32_t v[SCAN_UNROLL];
#pragma unroll
				for (int u = 0; u < SCAN_UNROLL; u++)
					v[u] = t + u < t1 ? blk[t + u] : 0u;
#pragma unroll
				for (int u = 0; u < SCAN_UNROLL; u++) {
					if (t + u < t1) blk[t + u] = run;
					run += v[u];
				}
			}
		}
	}
	if (total > P.cap) {
		if (threadIdx.x == 0) set_error(ctl, PP_OUT_OF_RESOURCES, tick, -1);
		return;
	}
	/* The replicas, in table order, by the first wave (no barriers). */
	if (threadIdx.x >= WAVE) return;
	const int lane = threadIdx.x;
	const uint32_t nrep = H.nrep;
	/* agent-steps, occupied, eaten, births */
	unsigned long long bs[4] = {0, 0, 0, 0};
	uint32_t carry = 0;
	for (uint32_t k0 = 0; k0 < nrep; k0 += WAVE) {
		const uint32_t k = k0 + lane;
		uint32_t slot = 0, out = 0;
		unsigned long long a[NACC];
		if (k < nrep) {
			slot = T.order[k];
#pragma unroll
			for (int i = 0; i < NACC; i++) a[i] = T.rctl[slot].acc[i];
			out = (uint32_t)a[AC_OUT];
		}
		const uint32_t x = wave_incl_scan32(out);
		const uint32_t pre = x - out;
		const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)x, WAVE - 1);
		if (k < nrep) {
			RepCtl &rc = T.rctl[slot];
			const Rep &rp = T.rep[slot];
			rc.base = carry + pre;
			if (rc.status == RS_ACTIVE) {
				const uint32_t s = tick - rp.t0;
				if (hi32(a[AC_OCC]) || out > rp.cap) {
					rc.status = RS_FAILED;
					rc.error = hi32(a[AC_OCC]) ? PP_ENERGY_OVERFLOW : PP_OUT_OF_RESOURCES;
					rc.err_step = s;
					rc.n = 0;
				} else {
					const uint32_t M = rp.R + 1;
					unsigned long long D = rc.dead_cells, S = rc.cd_sum;
					S -= D;
					const uint32_t expire = (tick + 1) % M;
					D -= rp.ring[expire];
					rp.ring[expire] = 0;
					const unsigned long long E = a[AC_EATEN];
					rp.ring[tick % M] += E;
					D += E;
					S += E * (unsigned long long)rp.R;
					rc.dead_cells = D;
					rc.cd_sum = S;
					write_row(a, rc, rp, &rp.stats[(s + 1) % rp.stats_rows]);
					bs[0] += rc.n;
					bs[1] += lo32(a[AC_OCC]);
					bs[2] += E;
					bs[3] += hi32(a[AC_OUT]);
					rc.agent_steps += rc.n;
					rc.step = s + 1;
					rc.n = out;
					if (s + 1 >= rp.iters) {
						rc.status = RS_DONE;
						rc.n = 0;
					}
				}
			}
#pragma unroll
			for (int i = 0; i < NACC; i++) rc.acc[i] = 0;
		}
		carry += tot;
	}
#pragma unroll
	for (int i = 0; i < 4; i++) bs[i] = wave_sum(bs[i]);
	if (lane != 0) return;
	ctl->agent_steps += bs[0];
	ctl->sum_live += L;
	ctl->sum_out += total;
	ctl->sum_occ += bs[1];
	ctl->sum_eaten += bs[2];
	ctl->sum_births += bs[3];
	ctl->n = total;
	ctl->nmixed = 0;
	ctl->live[(tick + 1) & 1] = 0;
	ctl->tick = tick + 1;
}

constexpr uint32_t BWIN = 128;   /* buckets per LDS count window */

/* Writes the staged outputs [0, total) of an action tile (batch positions
 * goff + i) to the SoA state (when keep: at the last tick of a run, since
 * no kernel reads it), moves each for the next iteration, and counts
 * each output in the bucket of its cell (ocnt: the order the next sort reads
 * them in) and each survivor of the move in the bucket of its new cell
 * (ncnt), through LDS windows of BWIN buckets from w0. MIXED = false: every
 * output belongs to replica m_u (its constants stay uniform across the
 * group); true: each looks up its own. Returns the survivors of the move. */
template <bool MIXED, int NT>
__device__ __forceinline__ unsigned long long emit_outputs(
	const uint32_t *s_oc, const uint32_t *s_ov, uint32_t *s_win,
	uint32_t *s_l1, uint32_t w0, uint32_t u0,
	uint32_t total, uint32_t goff, uint32_t *__restrict__ ocell,
	int32_t *__restrict__ oenergy, uint8_t *__restrict__ otype,
	uint32_t *__restrict__ mkeys, uint32_t *__restrict__ mvals,
	uint32_t *ocnt, uint32_t *ncnt, const Tables &T, uint32_t nrep,
	uint32_t tick, const RM &m_u, const Params &P, bool keep) {
	unsigned long long live = 0;
	const uint32_t S = P.bshift;
	for (uint32_t i0 = 0; i0 < total; i0 += NT) {
		const uint32_t i = i0 + threadIdx.x;
		uint32_t ob = NO_DIGIT, nbk = NO_DIGIT;
		if (i < total) {
			const uint32_t id = goff + i;
			const uint32_t cc = s_oc[i], ov = s_ov[i];
			const uint32_t en = ov & V_ENERGY, t = ov >> 31;
			if (keep) {
				at32(ocell, id) = cc;
				at32(oenergy, id) = (int32_t)en;
				at32(otype, id) = (uint8_t)t;
			}
			uint32_t mv = 0, mk = DEAD_KEY;
			const RM m = MIXED ? load_rm(T, rep_index(T, nrep, cc), tick, S) : m_u;
			const uint32_t l = cc - m.cb;
			const uint32_t y = div_by(l, m.gxm, m.gxs), x = l - y * m.gx;
			if (m.active)
				mk = move_agent(id - m.base, x, y, (int32_t)en, t, m, S, &mv, &nbk);
			ob = bucket_of(m.bbase, m.nb, m.h, x, y, S);
			if (P.bucket_on && m.nb >= 2) {
				const uint32_t xb = x & ((1u << S) - 1u);
				if (xb == 0) atomicAdd(&ocnt[edge_first_off(P.nb_cap) + ob], 1u);
				if (x + 1 == min(((x >> S) + 1) << S, m.gx))
					atomicAdd(&ocnt[edge_last_off(P.nb_cap) + ob], 1u);
			}
			__builtin_nontemporal_store(mk, &at32(mkeys, id));
			__builtin_nontemporal_store(mv, &at32(mvals, id));
			live += mk != DEAD_KEY;
		}
		if (P.bucket_on) {
			wave_bucket_add<BWIN>(s_win, s_l1, ocnt, P.nb_cap, P.cnt_top, w0, u0,
				ob);
			wave_bucket_add<BWIN>(s_win + BWIN, s_l1 + L1W, ncnt, P.nb_cap,
				P.cnt_top, w0, u0, nbk);
		}
	}
	return live;
}

/* The end of tick t for a batch of one replica (every plain simulation), by
 * the first wave of the emit's work-group 0: what k_act_scan does for a
 * batch of several, from the action counters the actions left in the
 * partial-sum slots (words of tick t's parity). The emit's other work-
 * groups read nothing it writes: they take the tick from Ctl::emit_tick and
 * derive the replica's state after the tick from the same counters. The
 * caller has checked the outputs against the agent buffers. */
__device__ void end_tick_solo(Ctl *ctl, const Tables &T, const Params &P,
	uint32_t tick, uint32_t L) {
	const int lane = threadIdx.x;
	const unsigned long long *q = P.part + lane * PART_STRIDE + part_acc(tick);
	unsigned long long a[NACC];
#pragma unroll
	for (int i = 0; i < NACC; i++) a[i] = wave_sum(q[i]);
	if (lane != 0) return;
	RepCtl &rc = T.rctl[P.one_slot];
	const Rep &rp = T.rep[P.one_slot];
	const uint32_t out = (uint32_t)a[AC_OUT];
	unsigned long long steps = 0, occ = 0, eaten = 0, births = 0;
	rc.base = 0;
	if (rc.status == RS_ACTIVE) {
		const uint32_t s = tick - rp.t0;
		if (hi32(a[AC_OCC]) || out > rp.cap) {
			rc.status = RS_FAILED;
			rc.error = hi32(a[AC_OCC]) ? PP_ENERGY_OVERFLOW : PP_OUT_OF_RESOURCES;
			rc.err_step = s;
			rc.n = 0;
		} else {
			const uint32_t M = rp.R + 1;
			unsigned long long D = rc.dead_cells, S = rc.cd_sum;
			S -= D;
			const uint32_t expire = (tick + 1) % M;
			D -= rp.ring[expire];
			rp.ring[expire] = 0;
			const unsigned long long E = a[AC_EATEN];
			rp.ring[tick % M] += E;
			D += E;
			S += E * (unsigned long long)rp.R;
			rc.dead_cells = D;
			rc.cd_sum = S;
			write_row(a, (unsigned long long)rp.C - D, S,
				&rp.stats[(s + 1) % rp.stats_rows]);
			steps = rc.n;
			occ = lo32(a[AC_OCC]);
			eaten = E;
			births = hi32(a[AC_OUT]);
			rc.agent_steps += rc.n;
			rc.step = s + 1;
			rc.n = out;
			if (s + 1 >= rp.iters) {
				rc.status = RS_DONE;
				rc.n = 0;
			}
		}
	}
	ctl->agent_steps += steps;
	ctl->sum_live += L;
	ctl->sum_out += out;
	ctl->sum_occ += occ;
	ctl->sum_eaten += eaten;
	ctl->sum_births += births;
	ctl->n = out;
	ctl->live[(tick + 1) & 1] = 0;
	ctl->tick = tick + 1;
}

/* Emits the new state and moves it for the next iteration. Sorted position
 * p emits its agent unless eaten, then its newborn if any; the agents' batch
 * positions in the new state are these outputs in position order (tile
 * offset from k_act_scan, or from the unscanned tile counts with one replica,
 * plus a work-group scan), and an agent's local id is
 * its position minus its replica's base. Each output is staged in LDS, then
 * written with coalesced stores to the SoA state and immediately moved for
 * tick t+1; the bucket counts of the next sort are summed on the way.
 * Agents of a replica that is done or failed get DEAD_KEY. */
template <int NT>
__global__ __launch_bounds__(NT) void k_emit_move(
	const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
	const uint32_t *__restrict__ act_off, uint32_t *__restrict__ ocell,
	int32_t *__restrict__ oenergy, uint8_t *__restrict__ otype,
	uint32_t *__restrict__ mkeys, uint32_t *__restrict__ mvals,
	uint32_t *ocnt, uint32_t *ncnt, Ctl *ctl, Tables T, Params P) {
	/* NT threads, IPT sorted positions each (the action tile) */
	constexpr int IPT = A_TILE / NT, NW = NT / WAVE;
	static_assert(IPT * NT == A_TILE && (IPT == 2 || IPT == 4), "emit shape");
	/* staged outputs: a tile's, after fewer than NT left from the tile
	 * before (one replica) */
	__shared__ uint32_t s_oc[A_STAGE + NT];
	__shared__ uint32_t s_l1[2 * L1W];
	__shared__ uint32_t s_l2[2 * L2W];
	__shared__ uint32_t s_ov[A_STAGE + NT];
	__shared__ uint32_t s_win[2 * BWIN];
	__shared__ uint32_t s_tmp[NW];
	__shared__ unsigned long long s_red[NW];

	const CtlHead H = ctl_head(ctl);
	const uint32_t state_tick = ctl->state_tick;
	const uint32_t emit_tick = ctl->emit_tick;
	if (H.error) return;
	const int tid = threadIdx.x;
	const bool solo = P.solo != 0;
	/* the tick being moved for (the acted tick + 1) */
	const uint32_t tick = solo ? emit_tick : H.tick;
	const bool keep = tick == state_tick;
	const uint32_t L = H.live(tick - 1);
	const uint32_t nrep = H.nrep;
	const uint32_t ntiles = (L + A_TILE - 1) / A_TILE;
	const uint32_t S = P.bshift;
	unsigned long long live = 0;
	/* A batch of one replica (every plain simulation) needs no lookups; one
	 * of several replicas reads its replica's constants per tile. */
	RM m_one{};
	if (solo) {
		/* The end of the acted tick without k_act_scan: every group sums
		 * the replica's outputs and energy-limit count from the partial-sum
		 * slots and derives the replica's state after the tick (the status
		 * it reads is the one before or, once group 0 has written it, after
		 * the tick: both give the same activity); group 0 also ends the
		 * tick. Outputs beyond the agent buffers fail the batch. */
		if (ntiles <= blockIdx.x && blockIdx.x != 0) return;
		const unsigned long long *q = P.part + (tid & (WAVE - 1)) * PART_STRIDE
			+ part_acc(tick - 1);
		/* the outputs (low halves: their sum stays below 2^32) and whether
		 * any slot counted an agent above the energy limit (high halves of
		 * the occupancy words: the occupied cells, in the low halves, stay
		 * below 2^32 in all, so no carry reaches them) */
		const uint32_t out = wave_sum32(lo32(q[AC_OUT]));
		const bool over = __ballot(hi32(q[AC_OCC]) != 0) != 0;
		if (out > P.cap) {
			if (blockIdx.x == 0 && tid == 0)
				set_error(ctl, PP_OUT_OF_RESOURCES, tick - 1, -1);
			return;
		}
		m_one = load_rm_uniform(T, 0, tick, S, P.one_slot);
		const Rep &rp = T.rep[P.one_slot];
		const uint32_t cap = __builtin_amdgcn_readfirstlane(rp.cap);
		const uint32_t t0 = __builtin_amdgcn_readfirstlane(rp.t0);
		const uint32_t iters = __builtin_amdgcn_readfirstlane(rp.iters);
		const bool failed = over || out > cap;
		m_one.active = m_one.active && !failed && tick - t0 < iters;
		m_one.base = 0;
		if (blockIdx.x == 0 && tid < WAVE) end_tick_solo(ctl, T, P, tick - 1, L);
	}

	for (uint32_t i = tid; i < 2 * BWIN; i += NT) s_win[i] = 0;
	for (uint32_t i = tid; i < 2 * L1W; i += NT) s_l1[i] = 0;
	if (tid < 2 * L2W) s_l2[tid] = 0;
	/* Flushes the count windows (bases w0, u0): level 0; level 1, summed
	 * into the level-2 window; levels 2 and 3. The windows persist over a
	 * group's run of tiles while its buckets stay inside them. */
	uint32_t w0 = 0, u0 = 0;
	bool have_win = false;
	auto flush = [&]() {
		for (uint32_t i = tid; i < 2 * BWIN; i += NT) {
			const uint32_t h = s_win[i];
			if (h) {
				atomicAdd(&(i < BWIN ? ocnt : ncnt)[w0 + (i % BWIN)], h);
				s_win[i] = 0;
			}
		}
		if (P.cnt_top >= 1) {
			for (uint32_t i = tid; i < 2 * L1W; i += NT) {
				const uint32_t h = s_l1[i];
				if (h) {
					const uint32_t k = i / L1W, q1 = u0 + (i % L1W);
					atomicAdd(&(k ? ncnt : ocnt)[cnt_off(P.nb_cap, 1) + q1], h);
					if (P.cnt_top >= 2)
						atomicAdd(&s_l2[k * L2W + (q1 >> 6) - (u0 >> 6)], h);
					s_l1[i] = 0;
				}
			}
			__syncthreads();
			if (P.cnt_top >= 2 && tid < 2 * L2W) {
				const uint32_t h = s_l2[tid];
				if (h) {
					const uint32_t k = tid / L2W, q2 = (u0 >> 6) + tid % L2W;
					uint32_t *cnt = k ? ncnt : ocnt;
					atomicAdd(&cnt[cnt_off(P.nb_cap, 2) + q2], h);
					s_l2[tid] = 0;
				}
			}
		}
		__syncthreads();
	};
	/* Each group takes a contiguous run of tiles: their cells, and so their
	 * buckets, ascend, and the windows are flushed rarely. */
	const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
	const uint32_t t_beg = blockIdx.x * per;
	const uint32_t t_end = min(ntiles, t_beg + per);
	/* One replica: the output offset of the group's first tile from the
	 * unscanned block sums (those before its block) and the counts of the
	 * tiles before it in its block; each later tile's adds the outputs of
	 * the one before. */
	uint32_t run_off = 0;
	if (solo && t_beg < t_end) {
		/* each wave on its own (no barrier) */
		const uint32_t nbl = t_beg >> 6, ln = (uint32_t)(tid & (WAVE - 1));
		uint32_t v = ln < (t_beg & 63u) ? act_off[(t_beg & ~63u) + ln] : 0u;
		for (uint32_t i = ln; i < nbl; i += WAVE) v += act_off[P.act_l1 + i];
		run_off = wave_sum32(v);
	}
	/* One replica: the outputs of the group's run of tiles have consecutive
	 * ids, so they are emitted as one stream in rounds of NT; the fewer than
	 * NT that do not fill a round wait in front of the next tile's (carry,
	 * ids from cbase), and only the run's last round is partial. */
	uint32_t carry = 0, cbase = 0;
	for (uint32_t tile = t_beg; tile < t_end; tile++) {
		const uint32_t base = tile * A_TILE;
		/* the tile's output offset: with several replicas, its block's
		 * scanned sum plus the counts of the block's tiles before it (a
		 * wave-wide load and sum, issued with the tile's pairs) */
		uint32_t goff = run_off;
		if (!solo) {
			const uint32_t tile_lane = tid & (WAVE - 1);
			const uint32_t before = tile_lane < (tile & 63u)
				? act_off[(tile & ~63u) + tile_lane] : 0u;
			goff = act_off[P.act_l1 + (tile >> 6)] + wave_sum32(before);
		}
		/* the tile's first cell, requested with its pairs (uniform: a
		 * scalar load) */
		const uint32_t c_first = __builtin_amdgcn_readfirstlane(keys[base]);
		const uint32_t p0 = base + tid * IPT;
		uint32_t c[IPT], v[IPT];
		/* (the last reads of the tick's pairs: non-temporal; 64 replicas
		 * 864 -> 858 us a tick, 12800 x 12800 632 -> 630 us) */
		if (p0 + IPT <= L) {
			if constexpr (IPT == 4) {
				typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
				const u32x4 kv = __builtin_nontemporal_load(&at32(reinterpret_cast<const u32x4 *>(keys), p0 / 4));
				const u32x4 vv = __builtin_nontemporal_load(&at32(reinterpret_cast<const u32x4 *>(vals), p0 / 4));
				c[0] = kv.x; c[1] = kv.y; c[2] = kv.z; c[3] = kv.w;
				v[0] = vv.x; v[1] = vv.y; v[2] = vv.z; v[3] = vv.w;
			} else {
				typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
				const u32x2 kv = __builtin_nontemporal_load(&at32(reinterpret_cast<const u32x2 *>(keys), p0 / 2));
				const u32x2 vv = __builtin_nontemporal_load(&at32(reinterpret_cast<const u32x2 *>(vals), p0 / 2));
				c[0] = kv.x; c[1] = kv.y;
				v[0] = vv.x; v[1] = vv.y;
			}
		} else {
#pragma unroll
			for (int j = 0; j < IPT; j++) {
				const bool in = p0 + j < L;
				c[j] = in ? keys[p0 + j] : 0u;
				v[j] = in ? vals[p0 + j] : V_KILLED;
			}
		}
		uint32_t nout = 0;
#pragma unroll
		for (int j = 0; j < IPT; j++)
			nout += ((v[j] & V_KILLED) == 0) + ((v[j] & V_CHILD) != 0);
		uint32_t total;
		uint32_t o = carry + wg_excl_scan<NT>(nout, s_tmp, total);
		if (carry == 0) cbase = goff;
		run_off += total;
#pragma unroll
		for (int j = 0; j < IPT; j++) {
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
		/* Replicas of this tile's positions (keys ascend), and the first
		 * bucket of the count window: the bucket of the tile's first cell,
		 * less BWIN/4 for moves to the row above. Uniform over the group. */
		bool mixed = false;
		RM m_t = m_one;
		uint32_t k_first = 0;
		if (nrep != 1) {
			k_first = rep_index(T, nrep, c_first);
			const uint32_t k_last = rep_index(T, nrep,
				__builtin_amdgcn_readfirstlane(keys[min(base + A_TILE, L) - 1]));
			mixed = k_first != k_last;
			m_t = load_rm_uniform(T, k_first, tick, S);
		}
		uint32_t nw0;
		{
			const uint32_t l = c_first - m_t.cb;
			const uint32_t y = l / m_t.gx, x = l - y * m_t.gx;
			const uint32_t b0 = bucket_of(m_t.bbase, m_t.nb, m_t.h, x, y, S);
			nw0 = b0 > BWIN / 4 ? b0 - BWIN / 4 : 0u;
		}
		/* re-centre the windows once the tiles' buckets have moved a
		 * quarter window up (buckets ascend over the run) */
		if (P.bucket_on && have_win && nw0 > w0 + BWIN / 4) {
			flush();
			have_win = false;
		}
		if (!have_win) {
			w0 = nw0;
			u0 = (w0 >> 6) > L1W / 2 ? (w0 >> 6) - L1W / 2 : 0u;
			have_win = true;
		}
		const uint32_t avail = carry + total;
		const uint32_t n_emit = solo && tile + 1 < t_end ? avail & ~(uint32_t)(NT - 1)
			: avail;
		if (!mixed)
			live += emit_outputs<false, NT>(s_oc, s_ov, s_win, s_l1, w0, u0, n_emit, cbase,
				ocell, oenergy, otype, mkeys, mvals, ocnt, ncnt, T, nrep, tick, m_t, P,
				keep);
		else
			live += emit_outputs<true, NT>(s_oc, s_ov, s_win, s_l1, w0, u0, n_emit, cbase,
				ocell, oenergy, otype, mkeys, mvals, ocnt, ncnt, T, nrep, tick, m_t, P,
				keep);
		__syncthreads();
		/* what did not fill a round moves to the front (from at least NT
		 * entries on: no overlap) */
		carry = avail - n_emit;
		cbase += n_emit;
		if (carry && n_emit) {
			if ((uint32_t)tid < carry) {
				s_oc[tid] = s_oc[n_emit + tid];
				s_ov[tid] = s_ov[n_emit + tid];
			}
			__syncthreads();
		}
	}
	if (P.bucket_on && have_win) flush();
	const int lane = tid & (WAVE - 1), w = tid / WAVE;
	/* a thread's survivors stay far below 2^32 (its share of one tick) */
	const unsigned long long tsum = wave_sum32((uint32_t)live);
	if (lane == 0) s_red[w] = tsum;
	__syncthreads();
	/* the next tick's first kernel adds up the slots into live[tick & 1] */
	if (tid == 0) {
		unsigned long long sum = 0;
		for (int i = 0; i < NW; i++) sum += s_red[i];
		if (sum)
			atomicAdd(&P.part[(blockIdx.x % PART_SLOTS) * PART_STRIDE + PART_LIVE],
				sum);
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
	KI_BUCKET_SCAN,
	KI_BUCKET_SORT,
	KI_COUNT
};

const char *const kernel_names[KI_COUNT] = {
	"init_reset", "init_cells", "init_agents", "init_done", "move0",
	"sort_up", "sort_scan", "sort_down", "act", "act_scan", "emit_move",
	"bucket_scan", "bucket_sort",
};

struct PendingEvent {
	int kernel;
	hipEvent_t start, stop;
};

/* Launch shape of the sort: keys per thread of a sort tile (tile = 256 x
 * kpt positions) and the widest digit (passes = ceil(key bits / max_bits),
 * bits spread evenly over the passes). */
struct SortShape {
	int kpt;
	int max_bits;
};

/* Host record of a replica. */
struct Replica {
	uint32_t id;
	int slot;              /* device slot, -1 once released */
	uint32_t cb, C;
	pp_params_t prm;
	uint64_t seed;
	uint32_t iters, reserve, n0;
	uint32_t status;       /* RS_* */
	int error;
	uint32_t err_step;
	uint32_t rows_have;    /* rows 0 .. rows_have - 1 collected */
	uint32_t n;            /* agents of its current state */
	uint64_t agent_steps;
	int keep;              /* keep every row on the host */
	std::vector<pp_stats64_t> rows;
	pp_stats64_t *sink;    /* else rows >= sink_from are written here */
	uint32_t sink_from;
};

uint32_t ceil_div(uint64_t a, uint64_t b) {
	return (uint32_t)((a + b - 1) / b);
}

uint32_t min_u32(uint32_t a, uint32_t b) { return a < b ? a : b; }

uint32_t grass_bytes_for(uint32_t R) {
	return R + REFRESH + 1 <= 256 ? 1 : R + REFRESH + 1 <= 65536 ? 2 : 4;
}

uint32_t key_bits_for(uint64_t cells) {
	uint32_t bits = 0;
	while (bits < 32 && ((cells - 1) >> bits) != 0) bits++;
	return bits ? bits : 1;
}

void shape_params(const SortShape &sh, uint32_t key_bits, uint32_t cap,
	uint32_t *passes, uint32_t *digit_bits, uint32_t *hist_cols) {
	*passes = (key_bits + sh.max_bits - 1) / sh.max_bits;
	*digit_bits = (key_bits + *passes - 1) / *passes;
	*hist_cols = ceil_div(cap, (uint64_t)WG * sh.kpt);
}

uint64_t hist_bytes_for(const SortShape &sh, uint32_t key_bits, uint32_t cap) {
	uint32_t passes, bits, cols;
	shape_params(sh, key_bits, cap, &passes, &bits, &cols);
	(void)passes;
	return (4ull << bits) * cols;
}

/* The default shape: tiles of 4096 keys; up to 22 key bits in passes of at
 * most 11 bits (1600 x 1600 = 2 x 11), wider keys in passes of at most 8
 * bits (12800 x 12800 = 4 x 7), where longer runs per bucket keep the
 * scatter coalesced. */
SortShape default_shape(uint32_t key_bits) {
	SortShape sh;
	sh.kpt = 16;
	const char *kpt_env = getenv("PPHPC_SORT_KPT");
	if (kpt_env && atoi(kpt_env) == 8) sh.kpt = 8;
	sh.max_bits = key_bits <= 22 ? 11 : 8;
	return sh;
}

/* Shapes the tuner may try: both tile sizes, both digit limits, as long as
 * the histogram table stays within max(32 MiB, the default's). */
std::vector<SortShape> tune_candidates(uint32_t key_bits, uint32_t cap) {
	const SortShape def = default_shape(key_bits);
	const uint64_t limit = std::max<uint64_t>(32ull << 20,
		hist_bytes_for(def, key_bits, cap));
	std::vector<SortShape> out;
	out.push_back(def);
	const int kpts[2] = {16, 8};
	const int bits[2] = {11, 8};
	for (int a = 0; a < 2; a++)
		for (int b = 0; b < 2; b++) {
			SortShape sh{kpts[a], bits[b]};
			if (sh.kpt == def.kpt && sh.max_bits == def.max_bits) continue;
			if (hist_bytes_for(sh, key_bits, cap) <= limit) out.push_back(sh);
		}
	return out;
}

} /* namespace */

struct pp_batch {
	int device = -1;
	int num_cu = 256;
	hipStream_t stream = nullptr;

	Params P{};
	uint32_t key_bits = 1;
	uint32_t max_rep = 0;
	uint32_t stats_rows = 0;
	uint32_t ring_stride = 0;
	uint32_t grass_bytes = 4;
	SortShape shape{16, 11};
	int rbits = 11;
	uint32_t grid_up = 0, grid_down = 0, grid_act = 0, grid_emit = 0;
	/* launch shape: rounds of resident groups of the action and emit grids,
	 * eighths of a round of the bucket sort's */
	uint32_t tile_rounds = TILE_ROUNDS, bucket_eighths = BUCKET_EIGHTHS;
	/* action and emit groups of WIDE_WG threads instead of WG (chosen from
	 * the live agents at each replica set) */
	int wide = 0;

	uint32_t *cell = nullptr;
	int32_t *energy = nullptr;
	uint8_t *type = nullptr;
	uint32_t *keys[3] = {nullptr, nullptr, nullptr};
	uint32_t *vals[3] = {nullptr, nullptr, nullptr};
	void *grass_at = nullptr;
	unsigned long long *ring = nullptr;
	uint32_t *hist = nullptr;
	uint64_t hist_alloc_words = 0;
	uint32_t hist_words = 0;
	uint32_t *totals = nullptr;
	uint32_t *act_cnt = nullptr;
	uint32_t *mixed_list = nullptr;  /* action tiles of several replicas */
	uint32_t act_words = 0;
	/* bucket sort: per-bucket counts by old and new cell (emit), all
	 * levels; first bucket of each replica of the table */
	uint32_t *ocnt = nullptr, *ncnt = nullptr;
	uint32_t *bbase_dev = nullptr;
	/* table index of each bucket's replica: nb_cap words after the
	 * survivor counts, in their allocation (the bucket sort reads it
	 * through ncnt, with no argument of its own) */
	uint32_t *brep_dev = nullptr;
	int need_radix = 1;   /* the next tick sorts with the radix sort */
	/* 0: radix sort every tick; 1: automatic; 8..11: row buckets of
	 * 2^mode cells */
	uint32_t sort_mode = 1;
	uint32_t grid_bucket = 0;
	/* the bucket sort reads records that k_bprefix (grid_bpre groups) and
	 * k_brec (grid_brec) write into the old-cell table's allocation */
	int bpre_on = 0;
	uint32_t grid_bpre = 0, grid_brec = 0;
	Ctl *ctl = nullptr;
	Ctl *ctl_host = nullptr;        /* pinned mirror */
	int ctl_fresh = 0;              /* mirror matches the device */
	Rep *rep_dev = nullptr;
	std::vector<Rep> rep_host;
	RepCtl *rctl_dev = nullptr;
	RepCtl *rctl_host = nullptr;    /* pinned mirror */
	uint32_t *cbs_dev = nullptr;
	uint32_t *crep_dev = nullptr;   /* Tables::crep, ceil(C / 2^CREP_SHIFT) words */
	uint32_t *order_dev = nullptr;
	pp_stats64_t *stats_dev = nullptr;
	pp_stats64_t *stats_host = nullptr; /* pinned, same layout */

	hipGraphExec_t graph_exec = nullptr;   /* 8 bucket-sort ticks */
	hipGraphExec_t graph_radix = nullptr;  /* 8 radix-sort ticks */
	int graph_steps = 0;

	int profiling = 0;
	int no_graph = 0;   /* never capture graphs (allocation-free runs) */
	std::vector<PendingEvent> pending;
	std::vector<hipEvent_t> free_events;
	uint64_t prof_launches[KI_COUNT] = {};
	double prof_ns[KI_COUNT] = {};

	std::vector<int64_t> slot_rid;            /* replica id per slot, -1 free */
	std::map<uint32_t, uint32_t> free_cells;  /* first cell -> length */
	uint64_t reserved = 0;
	std::map<uint32_t, Replica> reps;
	uint32_t next_id = 0;
	uint64_t table_version = 0;   /* replica table uploads */

	Tables tables() const {
		return Tables{rep_dev, rctl_dev, cbs_dev, order_dev, bbase_dev, crep_dev};
	}
};

struct pp_sim {
	pp_batch_t *b = nullptr;
	pp_params_t prm{};
	uint64_t seed = 0;
	uint32_t cap = 0;
	uint32_t rid = 0;
	int initialised = 0;
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

#define PP_TRY(expr)                                                        \
	do {                                                                    \
		int rc_ = (expr);                                                   \
		if (rc_ != PP_SUCCESS) return rc_;                                  \
	} while (0)

static hipEvent_t event_get(pp_batch_t *b) {
	if (!b->free_events.empty()) {
		hipEvent_t e = b->free_events.back();
		b->free_events.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	if (hipEventCreate(&e) != hipSuccess) return nullptr;
	return e;
}

/* Profiling wrappers: bracket a launch with events on the batch's stream. */
static void prof_begin(pp_batch_t *b, int kernel, PendingEvent &pe) {
	pe.kernel = kernel;
	pe.start = pe.stop = nullptr;
	if (!b->profiling) return;
	pe.start = event_get(b);
	pe.stop = event_get(b);
	if (pe.start) (void)hipEventRecord(pe.start, b->stream);
}

static void prof_end(pp_batch_t *b, PendingEvent &pe) {
	if (!b->profiling || !pe.start || !pe.stop) return;
	(void)hipEventRecord(pe.stop, b->stream);
	b->pending.push_back(pe);
}

static int prof_collect(pp_batch_t *b) {
	for (auto &pe : b->pending) {
		float ms = 0.f;
		HIP_OK(hipEventSynchronize(pe.stop));
		HIP_OK(hipEventElapsedTime(&ms, pe.start, pe.stop));
		b->prof_ns[pe.kernel] += 1e6 * (double)ms;
		b->prof_launches[pe.kernel] += 1;
		b->free_events.push_back(pe.start);
		b->free_events.push_back(pe.stop);
	}
	b->pending.clear();
	return PP_SUCCESS;
}

/* Pair buffers: the move writes buffer 0; pass p reads buffer 0 (p = 0) or
 * 1 + ((p - 1) & 1) and writes 1 + (p & 1). */
static int pass_src(uint32_t p) { return p == 0 ? 0 : 1 + (int)((p - 1) & 1); }
static int pass_dst(uint32_t p) { return 1 + (int)(p & 1); }

template <typename GT>
static void launch_sort_scan(pp_batch_t *b, uint32_t pass) {
	k_sort_scan<GT><<<1u << b->P.digit_bits, WG, 0, b->stream>>>(b->hist,
		b->totals, b->ctl, pass, b->P, b->act_cnt, b->act_words,
		(GT *)b->grass_at);
}

template <int KPT, int RBITS>
static void launch_sort(pp_batch_t *b, uint32_t pass) {
	PendingEvent pe;
	const int src = pass_src(pass), dst = pass_dst(pass);
	prof_begin(b, KI_SORT_UP, pe);
	k_sort_up<KPT><<<b->grid_up, WG, 0, b->stream>>>(b->keys[src], b->hist,
		b->ctl, pass, b->P);
	prof_end(b, pe);
	prof_begin(b, KI_SORT_SCAN, pe);
	if (b->grass_bytes == 1) launch_sort_scan<uint8_t>(b, pass);
	else if (b->grass_bytes == 2) launch_sort_scan<uint16_t>(b, pass);
	else launch_sort_scan<uint32_t>(b, pass);
	prof_end(b, pe);
	prof_begin(b, KI_SORT_DOWN, pe);
	k_sort_down<KPT, RBITS><<<b->grid_down, WG, 0, b->stream>>>(b->keys[src],
		b->vals[src], b->keys[dst], b->vals[dst], b->hist, b->totals, b->ctl,
		pass, b->P);
	prof_end(b, pe);
}

static void launch_sort_pass(pp_batch_t *b, uint32_t pass) {
	if (b->shape.kpt == 8) {
		if (b->rbits == 11) launch_sort<8, 11>(b, pass);
		else launch_sort<8, 8>(b, pass);
	} else {
		if (b->rbits == 11) launch_sort<16, 11>(b, pass);
		else launch_sort<16, 8>(b, pass);
	}
}

template <typename GT>
static void launch_act(pp_batch_t *b, int sorted) {
	if (b->wide)
		k_act<GT, WIDE_WG><<<b->grid_act, WIDE_WG, 0, b->stream>>>(b->keys[sorted],
			b->vals[sorted], (GT *)b->grass_at, b->act_cnt, b->mixed_list, b->ocnt,
			b->ncnt, b->ctl, b->tables(), b->P);
	else
		k_act<GT, WG><<<b->grid_act, WG, 0, b->stream>>>(b->keys[sorted],
			b->vals[sorted], (GT *)b->grass_at, b->act_cnt, b->mixed_list, b->ocnt,
			b->ncnt, b->ctl, b->tables(), b->P);
	/* A batch of one replica never has a tile of several. */
	if (b->max_rep > 1)
		k_act_mixed<GT><<<min_u32(b->max_rep, 256), WG, 0, b->stream>>>(
			b->keys[sorted], b->vals[sorted], (GT *)b->grass_at, b->act_cnt,
			b->mixed_list, b->ctl, b->tables(), b->P);
}

template <typename GT>
static void launch_bucket_sort(pp_batch_t *b) {
	if (b->bpre_on)
		k_bucket_sort<GT, true><<<b->grid_bucket, WG, 0, b->stream>>>(b->keys[0],
			b->vals[0], b->ocnt, b->ncnt, b->keys[1], b->vals[1],
			b->ctl, b->tables(), b->P, b->act_cnt, b->act_words, (GT *)b->grass_at);
	else
		k_bucket_sort<GT, false><<<b->grid_bucket, WG, 0, b->stream>>>(b->keys[0],
			b->vals[0], b->ocnt, b->ncnt, b->keys[1], b->vals[1],
			b->ctl, b->tables(), b->P, b->act_cnt, b->act_words, (GT *)b->grass_at);
}

/* Enqueue one tick (no synchronisation): the sort of the pairs the previous
 * tick's emit (or the moves at admission) produced -- the bucket sort, or the
 * radix sort when `radix` -- the actions, the scan + end of tick, and the
 * emit + move of the next tick. */
static int enqueue_step(pp_batch_t *b, int radix) {
	PendingEvent pe;
	int sorted = 1;
	if (radix && b->P.bucket_on) {
		/* counts of a previous layout are void */
		prof_begin(b, KI_BUCKET_SCAN, pe);
		k_bscan_clear<<<b->grid_bucket, WG, 0, b->stream>>>(b->ocnt, b->ncnt,
			cnt_words(b->P.nb_cap), b->ctl);
		prof_end(b, pe);
	}
	if (radix) {
		for (uint32_t pass = 0; pass < b->P.passes; pass++)
			launch_sort_pass(b, pass);
		sorted = pass_dst(b->P.passes - 1);
	} else {
		if (b->bpre_on) {
			prof_begin(b, KI_BUCKET_SCAN, pe);
			k_bprefix<<<b->grid_bpre, WG, 0, b->stream>>>(b->ocnt, b->ncnt, b->ocnt,
				b->ctl, b->P);
			k_brec<<<b->grid_brec, WG, 0, b->stream>>>(b->ocnt, b->ncnt, b->ctl,
				b->tables(), b->P);
			prof_end(b, pe);
		}
		prof_begin(b, KI_BUCKET_SORT, pe);
		if (b->grass_bytes == 1) launch_bucket_sort<uint8_t>(b);
		else if (b->grass_bytes == 2) launch_bucket_sort<uint16_t>(b);
		else launch_bucket_sort<uint32_t>(b);
		prof_end(b, pe);
	}
	prof_begin(b, KI_ACT, pe);
	if (b->grass_bytes == 1) launch_act<uint8_t>(b, sorted);
	else if (b->grass_bytes == 2) launch_act<uint16_t>(b, sorted);
	else launch_act<uint32_t>(b, sorted);
	prof_end(b, pe);
	/* the end of tick: k_act_scan for several replicas, the emit's work-
	 * group 0 for one */
	if (!b->P.solo) {
		prof_begin(b, KI_ACT_SCAN, pe);
		/* block sums that one wave scans in at most 8 steps: one wave */
		if (b->P.act_l1 <= 64u * 64u * 8u)
			k_act_scan<WAVE><<<1, WAVE, 0, b->stream>>>(b->act_cnt, b->ctl,
				b->tables(), b->P);
		else
			k_act_scan<SCAN_WG><<<1, SCAN_WG, 0, b->stream>>>(b->act_cnt, b->ctl,
				b->tables(), b->P);
		prof_end(b, pe);
	}
	prof_begin(b, KI_EMIT_MOVE, pe);
	if (b->wide)
		k_emit_move<WIDE_WG><<<b->grid_emit, WIDE_WG, 0, b->stream>>>(
			b->keys[sorted], b->vals[sorted], b->act_cnt, b->cell, b->energy,
			b->type, b->keys[0], b->vals[0], b->ocnt, b->ncnt, b->ctl,
			b->tables(), b->P);
	else
		k_emit_move<WG><<<b->grid_emit, WG, 0, b->stream>>>(b->keys[sorted],
			b->vals[sorted], b->act_cnt, b->cell, b->energy, b->type, b->keys[0],
			b->vals[0], b->ocnt, b->ncnt, b->ctl, b->tables(), b->P);
	prof_end(b, pe);
	HIP_OK(hipGetLastError());
	return PP_SUCCESS;
}

static void drop_graph(pp_batch_t *b) {
	if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
	if (b->graph_radix) (void)hipGraphExecDestroy(b->graph_radix);
	b->graph_exec = nullptr;
	b->graph_radix = nullptr;
	b->graph_steps = 0;
}

/* Work-groups of kernel f (WG threads) that the GPU holds at once. */
static uint32_t resident(const pp_batch_t *b, const void *f, int nt = WG) {
	int per_cu = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, nt, 0) !=
			hipSuccess || per_cu <= 0)
		per_cu = 2;
	return (uint32_t)per_cu * (uint32_t)b->num_cu;
}

template <typename GT>
static void set_grids_gt(pp_batch_t *b) {
	const Params &P = b->P;
	b->grid_act = min_u32(ceil_div(P.cap, A_TILE),
		b->tile_rounds * (b->wide
			? resident(b, (const void *)k_act<GT, WIDE_WG>, WIDE_WG)
			: resident(b, (const void *)k_act<GT, WG>)));
	b->grid_bucket = min_u32(P.nb_cap, std::max<uint32_t>(1u,
		b->bucket_eighths * resident(b, (const void *)k_bucket_sort<GT, false>) / 8));
}

/* Sets the sort shape and the launch grids that depend on it. */
static void apply_shape(pp_batch_t *b, const SortShape &sh) {
	drop_graph(b);
	b->shape = sh;
	Params &P = b->P;
	shape_params(sh, b->key_bits, P.cap, &P.passes, &P.digit_bits, &P.hist_cols);
	P.sort_tile = WG * sh.kpt;
	b->rbits = P.digit_bits > 8 ? 11 : 8;
	b->hist_words = (1u << P.digit_bits) * P.hist_cols;
	/* Enough groups to fill every CU several times, never more than the
	 * tiles of a full-capacity population; all loops are grid-strided. */
	const uint32_t cu = (uint32_t)b->num_cu;
	b->grid_up = min_u32(P.hist_cols, 8 * cu);
	/* (8-bit digits: three groups a CU, as many as its LDS holds; 11-bit
	 * digits: one) */
	b->grid_down = min_u32(P.hist_cols, (b->rbits == 11 ? 1 : 3) * cu);
	b->grid_emit = min_u32(ceil_div(P.cap, A_TILE), b->tile_rounds *
		(b->wide ? resident(b, (const void *)k_emit_move<WIDE_WG>, WIDE_WG)
			: resident(b, (const void *)k_emit_move<WG>)));
	if (b->grass_bytes == 1) set_grids_gt<uint8_t>(b);
	else if (b->grass_bytes == 2) set_grids_gt<uint16_t>(b);
	else set_grids_gt<uint32_t>(b);
}

/* Wide action and emit groups while the tiles of `agents` fit in one round
 * of wide emit groups. */
static int wide_auto(const pp_batch_t *b, uint64_t agents) {
	const uint64_t tiles = (agents + A_TILE - 1) / A_TILE;
	return tiles <= resident(b, (const void *)k_emit_move<WIDE_WG>, WIDE_WG)
		? 1 : 0;
}

/* Sets the launch shape (grids of the action, emit and bucket-sort kernels). */
static void apply_launch(pp_batch_t *b, uint32_t tile_rounds,
	uint32_t bucket_eighths) {
	drop_graph(b);
	b->tile_rounds = tile_rounds;
	b->bucket_eighths = bucket_eighths;
	apply_shape(b, b->shape);
}

/* Capture `steps` ticks into one executable graph (kernels only; no memset
 * or copy nodes). */
static int build_graph(pp_batch_t *b, int steps, int radix,
	hipGraphExec_t *out) {
	hipGraph_t g = nullptr;
	HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
	int rc = PP_SUCCESS;
	for (int i = 0; i < steps && rc == PP_SUCCESS; i++) rc = enqueue_step(b, radix);
	hipError_t e = hipStreamEndCapture(b->stream, &g);
	if (rc != PP_SUCCESS) {
		if (g) (void)hipGraphDestroy(g);
		return rc;
	}
	HIP_OK(e);
	e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
	(void)hipGraphDestroy(g);
	HIP_OK(e);
	b->graph_steps = steps;
	return PP_SUCCESS;
}

/* Enqueue `ticks` ticks: after a change of the replica table the first one
 * sorts with the radix sort (the moved order is not the cell order yet, or
 * the bucket layout changed); then whole graphs of 8, then single launches.
 * The last tick's emit writes the SoA state that state_of reads. */
static int enqueue_ticks(pp_batch_t *b, uint32_t ticks) {
	b->ctl_fresh = 0;
	const int use_graph = !b->profiling && !b->no_graph &&
		getenv("PPHPC_NO_GRAPH") == nullptr;
	const int radix = !b->P.bucket_on;
	uint32_t i = 0;
	if (ticks) {
		k_state_tick<<<1, WAVE, 0, b->stream>>>(b->ctl, ticks);
		HIP_OK(hipGetLastError());
	}
	if (ticks && b->need_radix && !radix) {
		PP_TRY(enqueue_step(b, 1));
		i = 1;
	}
	b->need_radix = 0;
	if (use_graph && ticks - i >= 8) {
		hipGraphExec_t *g = radix ? &b->graph_radix : &b->graph_exec;
		if (!*g) PP_TRY(build_graph(b, 8, radix, g));
		for (; i + 8 <= ticks; i += 8) HIP_OK(hipGraphLaunch(*g, b->stream));
	}
	for (; i < ticks; i++) PP_TRY(enqueue_step(b, radix));
	return PP_SUCCESS;
}

static int read_ctl(pp_batch_t *b) {
	HIP_OK(hipMemcpyAsync(b->ctl_host, b->ctl, sizeof(Ctl),
		hipMemcpyDeviceToHost, b->stream));
	HIP_OK(hipMemcpyAsync(b->rctl_host, b->rctl_dev,
		(size_t)b->max_rep * sizeof(RepCtl), hipMemcpyDeviceToHost, b->stream));
	HIP_OK(hipStreamSynchronize(b->stream));
	b->ctl_fresh = 1;
	return prof_collect(b);
}

/* Replica table: the slots in use sorted by first cell, and its length. The
 * stream must be idle. */
static int upload_table(pp_batch_t *b) {
	std::vector<std::pair<uint32_t, uint32_t>> t;
	for (uint32_t s = 0; s < b->max_rep; s++)
		if (b->slot_rid[s] >= 0) t.push_back({b->rep_host[s].cb, s});
	std::sort(t.begin(), t.end());
	std::vector<uint32_t> cbs(t.size()), order(t.size());
	for (size_t i = 0; i < t.size(); i++) {
		cbs[i] = t[i].first;
		order[i] = t[i].second;
	}
	if (!t.empty()) {
		HIP_OK(hipMemcpy(b->cbs_dev, cbs.data(), t.size() * 4,
			hipMemcpyHostToDevice));
		HIP_OK(hipMemcpy(b->order_dev, order.data(), t.size() * 4,
			hipMemcpyHostToDevice));
	}
	const uint32_t nrep = (uint32_t)t.size();
	/* the cell-chunk table of rep_index, for a table of several */
	if (nrep > 1) {
		std::vector<uint32_t> crep((size_t)(b->P.C >> CREP_SHIFT) + 1);
		uint32_t k = 0;
		for (size_t i = 0; i < crep.size(); i++) {
			const uint64_t c0 = (uint64_t)i << CREP_SHIFT;
			while (k + 1 < nrep && cbs[k + 1] <= c0) k++;
			crep[i] = k;
		}
		HIP_OK(hipMemcpy(b->crep_dev, crep.data(), crep.size() * 4,
			hipMemcpyHostToDevice));
	}
	/* bucket layout: rows of blocks of 2^bshift cells, replicas in table
	 * order */
	const uint32_t shift = b->sort_mode >= 8 ? b->sort_mode : 11u;
	std::vector<uint32_t> bbase(nrep + 1);
	uint64_t nbk = 0;
	for (uint32_t i = 0; i < nrep; i++) {
		bbase[i] = (uint32_t)std::min<uint64_t>(nbk, 0xFFFFFFFFull);
		const Rep &rp = b->rep_host[order[i]];
		const uint32_t h = rows_per_bucket(rp.gx, shift);
		nbk += (uint64_t)((rp.gy + h - 1) / h) * row_blocks(rp.gx, shift);
	}
	bbase[nrep] = (uint32_t)std::min<uint64_t>(nbk, 0xFFFFFFFFull);
	HIP_OK(hipMemcpy(b->bbase_dev, bbase.data(), (nrep + 1) * 4,
		hipMemcpyHostToDevice));
	/* Row buckets when chosen, or automatically: always for one replica,
	 * for several while the buckets are few or dense enough that their
	 * per-bucket latency stays below the radix sort's passes. The density
	 * counts the replicas' agents as last collected (their initial agents
	 * before a first collection). */
	uint64_t agents = 0;
	for (const auto &kv : b->reps)
		if (kv.second.slot >= 0) agents += kv.second.n;
	const uint32_t on = (b->sort_mode != 0 && nbk <= b->P.nb_cap &&
		nbk <= CNT_MAX_BUCKETS &&
		(b->sort_mode != 1 || nrep <= 1 || nbk <= BUCKET_AUTO_MAX ||
			agents >= BUCKET_AUTO_DENSITY * nbk)) ? 1u : 0u;
	uint32_t top = 0;
	while (top + 1 < (uint32_t)CNT_LEVELS && (nbk >> (6 * (top + 1))) > 0) top++;
	/* the replica of each bucket, for the bucket sort of several */
	if (on && nrep > 1) {
		std::vector<uint32_t> brep(nbk);
		for (uint32_t i = 0; i < nrep; i++)
			std::fill(brep.begin() + bbase[i], brep.begin() + bbase[i + 1], i);
		HIP_OK(hipMemcpy(b->brep_dev, brep.data(), nbk * 4, hipMemcpyHostToDevice));
	}
	const uint32_t one = nrep == 1 ? order[0] : 0u;
	const uint32_t solo = nrep == 1 ? 1u : 0u;
	/* prefix tables for a large layout (PPHPC_BPREFIX_MIN sets the bucket
	 * count from which they are used; results do not depend on it) */
	uint64_t bpre_min = BPRE_MIN;
	if (const char *e = getenv("PPHPC_BPREFIX_MIN")) bpre_min = strtoull(e, nullptr, 10);
	const int bpre = on && nbk >= bpre_min ? 1 : 0;
	const uint32_t gpre = bpre ? ceil_div(2 * ((nbk >> 6) + 1), NWAVE) : 0u;
	const uint32_t grec = bpre ? ceil_div(nbk, WG) : 0u;
	if (bpre != b->bpre_on || gpre != b->grid_bpre || grec != b->grid_brec) {
		b->bpre_on = bpre;
		b->grid_bpre = gpre;
		b->grid_brec = grec;
		drop_graph(b);
	}
	/* group width: PPHPC_WIDE_GROUPS (0 or 1), or wide while the agents'
	 * tiles fit in one round of wide emit groups (results do not depend on
	 * it) */
	int wide = wide_auto(b, agents);
	if (const char *e = getenv("PPHPC_WIDE_GROUPS")) wide = atoi(e) ? 1 : 0;
	if (wide != b->wide) {
		b->wide = wide;
		apply_shape(b, b->shape);    /* the action and emit grids */
	}
	if (on != b->P.bucket_on || shift != b->P.bshift || top != b->P.cnt_top ||
		one != b->P.one_slot || solo != b->P.solo) {
		b->P.bucket_on = on;
		b->P.bshift = shift;
		b->P.cnt_top = top;
		b->P.one_slot = one;
		b->P.solo = solo;
		drop_graph(b);     /* Params are launch arguments of the graphs */
	}
	const uint32_t nb32 = on ? (uint32_t)nbk : 0u;
	HIP_OK(hipMemcpy(&b->ctl->nrep, &nrep, 4, hipMemcpyHostToDevice));
	HIP_OK(hipMemcpy(&b->ctl->nbuckets, &nb32, 4, hipMemcpyHostToDevice));
	b->ctl_host->nrep = nrep;
	b->ctl_host->nbuckets = nb32;
	b->need_radix = 1;
	b->table_version++;
	return PP_SUCCESS;
}

static void release_cells(pp_batch_t *b, uint32_t cb, uint32_t C) {
	auto it = b->free_cells.emplace(cb, C).first;
	auto next = std::next(it);
	if (next != b->free_cells.end() && it->first + it->second == next->first) {
		it->second += next->second;
		b->free_cells.erase(next);
	}
	if (it != b->free_cells.begin()) {
		auto prev = std::prev(it);
		if (prev->first + prev->second == it->first) {
			prev->second += it->second;
			b->free_cells.erase(it);
		}
	}
}

/* First fit in the free cell ranges; returns the first cell or -1. */
static int64_t take_cells(pp_batch_t *b, uint32_t C) {
	for (auto it = b->free_cells.begin(); it != b->free_cells.end(); ++it) {
		if (it->second < C) continue;
		const uint32_t cb = it->first, len = it->second;
		b->free_cells.erase(it);
		if (len > C) b->free_cells.emplace(cb + C, len - C);
		return cb;
	}
	return -1;
}

static void release_slot(pp_batch_t *b, Replica &r) {
	if (r.slot < 0) return;
	b->slot_rid[r.slot] = -1;
	release_cells(b, r.cb, r.C);
	b->reserved -= r.reserve;
	r.slot = -1;
}

/* After the stream is idle: new statistics rows of every replica (to its
 * host rows or sink), replica states, and the release of the slots, cells
 * and reservations of replicas that are done or failed. */
static int collect(pp_batch_t *b) {
	PP_TRY(read_ctl(b));
	if (b->ctl_host->error && getenv("PPHPC_DEBUG"))
		fprintf(stderr, "pphpc: batch error %d at tick %u (detail %d)\n",
			b->ctl_host->error, b->ctl_host->err_step, b->ctl_host->err_rep);
	std::vector<Replica *> got;
	for (auto &kv : b->reps) {
		Replica &r = kv.second;
		if (r.slot < 0) continue;
		const RepCtl &rc = b->rctl_host[r.slot];
		if (rc.status == RS_FREE) continue;
		const uint32_t last = rc.step;  /* rows 0..last exist */
		if (last + 1 > r.rows_have) {
			const uint32_t first = r.rows_have, nrows = last + 1 - first;
			const uint32_t R = b->stats_rows;
			pp_stats64_t *dev = b->stats_dev + (size_t)r.slot * R;
			pp_stats64_t *host = b->stats_host + (size_t)r.slot * R;
			const uint32_t a = first % R;
			const uint32_t n1 = min_u32(nrows, R - a);
			HIP_OK(hipMemcpyAsync(host + a, dev + a, (size_t)n1 * sizeof(pp_stats64_t),
				hipMemcpyDeviceToHost, b->stream));
			if (nrows > n1)
				HIP_OK(hipMemcpyAsync(host, dev, (size_t)(nrows - n1) *
					sizeof(pp_stats64_t), hipMemcpyDeviceToHost, b->stream));
			got.push_back(&r);
		}
	}
	if (!got.empty()) HIP_OK(hipStreamSynchronize(b->stream));
	int dirty = 0;
	for (auto &kv : b->reps) {
		Replica &r = kv.second;
		if (r.slot < 0) continue;
		const RepCtl &rc = b->rctl_host[r.slot];
		if (rc.status == RS_FREE) continue;
		const uint32_t R = b->stats_rows;
		const pp_stats64_t *host = b->stats_host + (size_t)r.slot * R;
		for (uint32_t i = r.rows_have; i <= rc.step; i++) {
			const pp_stats64_t &row = host[i % R];
			if (r.keep) {
				if (r.rows.size() <= i) r.rows.resize(i + 1);
				r.rows[i] = row;
			} else if (r.sink && i >= r.sink_from) {
				r.sink[i - r.sink_from] = row;
			}
		}
		r.rows_have = rc.step + 1;
		r.n = rc.n;
		r.agent_steps = rc.agent_steps;
		r.status = rc.status;
		if (rc.status == RS_FAILED) {
			r.error = rc.error;
			r.err_step = rc.err_step;
		}
		if (rc.status == RS_DONE || rc.status == RS_FAILED) {
			release_slot(b, r);
			dirty = 1;
		}
	}
	if (dirty) PP_TRY(upload_table(b));
	return PP_SUCCESS;
}

static uint32_t active_count(const pp_batch_t *b) {
	uint32_t n = 0;
	for (auto &kv : b->reps)
		if (kv.second.slot >= 0 && kv.second.status == RS_ACTIVE) n++;
	return n;
}

/* Runs up to `ticks` ticks in chunks that the stats rings hold, collecting
 * after each; stops early when no replica is active or on a batch error. */
static int advance(pp_batch_t *b, uint32_t ticks) {
	HIP_OK(hipSetDevice(b->device));
	if (!b->ctl_fresh) PP_TRY(read_ctl(b));
	if (b->ctl_host->error) return b->ctl_host->error;
	const uint64_t last = (uint64_t)b->ctl_host->tick + ticks + b->P.R + 1;
	if (last >= 0xFFFFFFFFull) return PP_INVALID_ARGS;
	uint32_t done = 0;
	const uint32_t chunk_max = b->stats_rows / 2;
	while (done < ticks && active_count(b) > 0) {
		const uint32_t chunk = min_u32(ticks - done, chunk_max);
		PP_TRY(enqueue_ticks(b, chunk));
		PP_TRY(collect(b));
		done += chunk;
		if (b->ctl_host->error) return b->ctl_host->error;
	}
	return PP_SUCCESS;
}

static void batch_free(pp_batch_t *b) {
	if (!b) return;
	if (b->device >= 0) (void)hipSetDevice(b->device);
	drop_graph(b);
	for (auto &pe : b->pending) {
		(void)hipEventDestroy(pe.start);
		(void)hipEventDestroy(pe.stop);
	}
	for (auto e : b->free_events) (void)hipEventDestroy(e);
	(void)hipFree(b->cell);
	(void)hipFree(b->energy);
	(void)hipFree(b->type);
	for (int i = 0; i < 3; i++) {
		(void)hipFree(b->keys[i]);
		(void)hipFree(b->vals[i]);
	}
	(void)hipFree(b->grass_at);
	(void)hipFree(b->ring);
	(void)hipFree(b->hist);
	(void)hipFree(b->totals);
	(void)hipFree(b->act_cnt);
	(void)hipFree(b->mixed_list);
	(void)hipFree(b->ocnt);
	(void)hipFree(b->ncnt);
	(void)hipFree(b->bbase_dev);
	(void)hipFree(b->ctl);
	(void)hipFree(b->P.part);
	(void)hipFree(b->rep_dev);
	(void)hipFree(b->rctl_dev);
	(void)hipFree(b->cbs_dev);
	(void)hipFree(b->crep_dev);
	(void)hipFree(b->order_dev);
	(void)hipFree(b->stats_dev);
	if (b->stats_host) (void)hipHostFree(b->stats_host);
	if (b->ctl_host) (void)hipHostFree(b->ctl_host);
	if (b->rctl_host) (void)hipHostFree(b->rctl_host);
	if (b->stream) (void)hipStreamDestroy(b->stream);
	delete b;
}

/* Device bytes of a batch (the allocations of batch_alloc). */
/* Bucket capacity of the count tables: one bucket per 64 cells and 4096
 * more (rows narrower than a block have one bucket each; a layout with more
 * buckets than this sorts with the radix sort). */
static uint32_t nb_cap_for(uint64_t cells) {
	return (uint32_t)std::min<uint64_t>(cells / 64 + 4096, 0x7FFFFFFFull);
}

static uint64_t batch_bytes(uint64_t A, uint64_t C, uint32_t max_rep,
	uint32_t R, uint32_t stats_rows, uint64_t hist_words) {
	const uint64_t A4 = (A + 3) & ~3ull;
	uint64_t t = A4 * (4 + 4 + 1) + 6 * A4 * 4;
	t += 8ull * cnt_words(nb_cap_for(C)) + 4ull * (max_rep + 1);
	t += 4ull * nb_cap_for(C);   /* the replica of each bucket */
	/* the prefix tables and bucket records */
	t += 4ull * (brec_off(nb_cap_for(C)) - cnt_words(nb_cap_for(C)))
		+ 4ull * BR_WORDS * nb_cap_for(C);
	t += 4ull * ((C >> CREP_SHIFT) + 1);   /* the cell-chunk table */
	t += C * grass_bytes_for(R);
	t += (uint64_t)max_rep * (R + 1) * 8;
	t += hist_words * 4 + RADIX_MAX * 4;
	{
		/* tile counts and their sums per 64 tiles; the mixed-tile list */
		const uint64_t tiles = (uint64_t)ceil_div(A, A_TILE) + 1;
		t += (tiles + (tiles + 63) / 64 + 1) * 4 + tiles * 4;
	}
	t += sizeof(Ctl) + PART_SLOTS * PART_STRIDE * 8 + (uint64_t)max_rep * (sizeof(Rep) + sizeof(RepCtl) + 8);
	t += (uint64_t)max_rep * stats_rows * sizeof(pp_stats64_t);
	return t;
}

static uint32_t stats_rows_for(uint32_t max_rep) {
	return max_rep == 1 ? STATS_RING : 1024;
}

static uint64_t max_hist_words(uint32_t key_bits, uint32_t cap) {
	uint64_t w = 0;
	for (const SortShape &sh : tune_candidates(key_bits, cap))
		w = std::max(w, hist_bytes_for(sh, key_bits, cap) / 4);
	return w;
}

/* The device buffers of a batch, each taken from take(&ptr, bytes): device
 * allocations (batch_alloc), or slices of the static scratch region
 * (pp_sim_simulate_noalloc). b->hist_alloc_words is set by the caller. */
template <typename Take>
static int batch_buffers(pp_batch_t *b, Take &&take) {
	const uint64_t A = b->P.cap, C = b->P.C;
	/* Vector accesses read/write 16 bytes at multiples of 4 agents. */
	const uint64_t A4 = (A + 3) & ~3ull;
	HIP_OK(take((void **)&b->cell, A4 * 4));
	HIP_OK(take((void **)&b->energy, A4 * 4));
	HIP_OK(take((void **)&b->type, A4));
	for (int i = 0; i < 3; i++) {
		HIP_OK(take((void **)&b->keys[i], A4 * 4));
		HIP_OK(take((void **)&b->vals[i], A4 * 4));
	}
	HIP_OK(take((void **)&b->grass_at, C * b->grass_bytes));
	b->ring_stride = b->P.R + 1;
	HIP_OK(take((void **)&b->ring, (size_t)b->max_rep * b->ring_stride * 8));
	HIP_OK(take((void **)&b->hist, (size_t)b->hist_alloc_words * 4));
	HIP_OK(take((void **)&b->totals, RADIX_MAX * 4));
	/* tile counts, then the sums of blocks of 64 tiles */
	const uint32_t act_tiles = ceil_div(A, A_TILE) + 1;
	b->P.act_l1 = act_tiles;
	b->act_words = act_tiles + ceil_div(act_tiles, 64) + 1;
	HIP_OK(take((void **)&b->act_cnt, (size_t)b->act_words * 4));
	HIP_OK(take((void **)&b->mixed_list, (size_t)act_tiles * 4));
	const size_t nbw = cnt_words(b->P.nb_cap);
	/* the old-cell table, then the prefix tables and bucket records */
	HIP_OK(take((void **)&b->ocnt,
		((size_t)brec_off(b->P.nb_cap) + (size_t)BR_WORDS * b->P.nb_cap) * 4));
	HIP_OK(take((void **)&b->ncnt, (nbw + b->P.nb_cap) * 4));
	b->brep_dev = b->ncnt + nbw;
	HIP_OK(take((void **)&b->bbase_dev, ((size_t)b->max_rep + 1) * 4));
	HIP_OK(take((void **)&b->ctl, sizeof(Ctl)));
	HIP_OK(take((void **)&b->P.part, PART_SLOTS * PART_STRIDE * 8));
	HIP_OK(take((void **)&b->rep_dev, (size_t)b->max_rep * sizeof(Rep)));
	HIP_OK(take((void **)&b->rctl_dev, (size_t)b->max_rep * sizeof(RepCtl)));
	HIP_OK(take((void **)&b->cbs_dev, (size_t)b->max_rep * 4));
	HIP_OK(take((void **)&b->crep_dev, ((size_t)(C >> CREP_SHIFT) + 1) * 4));
	HIP_OK(take((void **)&b->order_dev, (size_t)b->max_rep * 4));
	const size_t srows = (size_t)b->max_rep * b->stats_rows;
	HIP_OK(take((void **)&b->stats_dev, srows * sizeof(pp_stats64_t)));
	return PP_SUCCESS;
}

static int batch_alloc(pp_batch_t *b) {
	b->hist_alloc_words = max_hist_words(b->key_bits, b->P.cap);
	PP_TRY(batch_buffers(b, [](void **p, size_t bytes) {
		return hipMalloc(p, bytes);
	}));
	const size_t nbw = cnt_words(b->P.nb_cap);
	HIP_OK(hipMemset(b->ocnt, 0, nbw * 4));
	HIP_OK(hipMemset(b->ncnt, 0, nbw * 4));
	HIP_OK(hipMemset(b->P.part, 0, PART_SLOTS * PART_STRIDE * 8));
	const size_t srows = (size_t)b->max_rep * b->stats_rows;
	HIP_OK(hipHostMalloc(&b->stats_host, srows * sizeof(pp_stats64_t),
		hipHostMallocDefault));
	HIP_OK(hipHostMalloc(&b->ctl_host, sizeof(Ctl), hipHostMallocDefault));
	HIP_OK(hipHostMalloc(&b->rctl_host, (size_t)b->max_rep * sizeof(RepCtl),
		hipHostMallocDefault));
	HIP_OK(hipMemset(b->ctl, 0, sizeof(Ctl)));
	HIP_OK(hipMemset(b->rctl_dev, 0, (size_t)b->max_rep * sizeof(RepCtl)));
	HIP_OK(hipDeviceSynchronize());
	memset(b->ctl_host, 0, sizeof(Ctl));
	memset(b->rctl_host, 0, (size_t)b->max_rep * sizeof(RepCtl));
	b->rep_host.assign(b->max_rep, Rep{});
	b->slot_rid.assign(b->max_rep, -1);
	return PP_SUCCESS;
}

static int batch_reset(pp_batch_t *b);

static int batch_create(uint64_t agents, uint64_t cells, uint32_t max_rep,
	uint32_t max_R, int device, pp_batch_t **out) {
	*out = nullptr;
	/* Output offsets and ids are 32-bit, and an iteration can at most double
	 * the survivors. */
	if (agents == 0 || agents >= (1ull << 29)) return PP_INVALID_ARGS;
	if (cells == 0 || cells > 0xFFFFFFFEull) return PP_INVALID_ARGS;
	if (max_rep == 0 || max_rep > 65536) return PP_INVALID_ARGS;
	if (max_R == 0 || max_R > (1u << 24)) return PP_INVALID_ARGS;
	pp_batch_t *b = new pp_batch_t();
	b->device = device;
	if (device < 0 && hipGetDevice(&b->device) != hipSuccess) {
		delete b;
		return PP_LIBRARY_ERROR;
	}
	if (hipSetDevice(b->device) != hipSuccess) {
		b->device = -1;
		batch_free(b);
		return PP_LIBRARY_ERROR;
	}
	if (hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount,
			b->device) != hipSuccess || b->num_cu <= 0)
		b->num_cu = 256;
	if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
		batch_free(b);
		return PP_LIBRARY_ERROR;
	}
	b->P.cap = (uint32_t)agents;
	b->P.C = (uint32_t)cells;
	b->P.R = max_R;
	b->P.nb_cap = nb_cap_for(cells);
	b->P.bshift = 11;
	const char *bs_env = getenv("PPHPC_SORT");
	if (bs_env) {
		const int v = atoi(bs_env);
		if (v == 0 || v == 1 || (v >= 8 && v <= 11)) b->sort_mode = (uint32_t)v;
	}
	b->max_rep = max_rep;
	b->stats_rows = stats_rows_for(max_rep);
	b->key_bits = key_bits_for(cells);
	/* Narrowest grass storage whose modular window holds R + REFRESH + 1. */
	b->grass_bytes = grass_bytes_for(max_R);
	const char *gb_env = getenv("PPHPC_GRASS_BYTES");
	if (gb_env) {
		const int v = atoi(gb_env);
		if ((v == 2 && max_R + REFRESH + 1 <= 65536) || v == 4) b->grass_bytes = v;
	}
	apply_shape(b, default_shape(b->key_bits));
	int rc = batch_alloc(b);
	if (rc != PP_SUCCESS) {
		batch_free(b);
		return rc;
	}
	rc = batch_reset(b);
	if (rc != PP_SUCCESS) {
		batch_free(b);
		return rc;
	}
	*out = b;
	return PP_SUCCESS;
}

/* Empties the batch (tick 0, no replicas). */
static int batch_reset(pp_batch_t *b) {
	HIP_OK(hipSetDevice(b->device));
	HIP_OK(hipStreamSynchronize(b->stream));
	PendingEvent pe;
	prof_begin(b, KI_INIT_RESET, pe);
	k_batch_reset<<<min_u32(ceil_div(b->hist_alloc_words, WG), 2048), WG, 0,
		b->stream>>>(b->ctl, b->hist, (uint32_t)b->hist_alloc_words, b->P.part);
	prof_end(b, pe);
	HIP_OK(hipGetLastError());
	HIP_OK(hipMemsetAsync(b->rctl_dev, 0, (size_t)b->max_rep * sizeof(RepCtl),
		b->stream));
	HIP_OK(hipStreamSynchronize(b->stream));
	memset(b->ctl_host, 0, sizeof(Ctl));
	memset(b->rctl_host, 0, (size_t)b->max_rep * sizeof(RepCtl));
	b->ctl_fresh = 1;
	b->reps.clear();
	b->slot_rid.assign(b->max_rep, -1);
	b->free_cells.clear();
	b->free_cells[0] = b->P.C;
	b->reserved = 0;
	b->need_radix = 1;
	return PP_SUCCESS;
}

static int check_params(const pp_params_t *p) {
	if (!p) return PP_INVALID_ARGS;
	const uint64_t C = (uint64_t)p->grid_x * p->grid_y;
	if (p->grid_x == 0 || p->grid_y == 0 || C > 0xFFFFFFFEull)
		return PP_INVALID_ARGS;
	if (p->grass_restart == 0 || p->grass_restart > (1u << 24))
		return PP_INVALID_ARGS;
	if (p->sheep_gain_from_food > (1u << 20) || p->wolves_gain_from_food > (1u << 20))
		return PP_INVALID_ARGS;
	if ((uint64_t)p->init_sheep + p->init_wolves >= (1ull << 29))
		return PP_INVALID_ARGS;
	return PP_SUCCESS;
}

/* Admits one replica: its slot, cells and agent reservation, its initial
 * state appended to the batch state and moved for its first iteration. */
static int admit(pp_batch_t *b, const pp_params_t *p, uint64_t seed,
	uint64_t reserve, uint32_t iters, int keep, uint32_t *id_out) {
	PP_TRY(check_params(p));
	HIP_OK(hipSetDevice(b->device));
	const uint32_t C = p->grid_x * p->grid_y;
	const uint32_t n0 = p->init_sheep + p->init_wolves;
	if (p->grass_restart > b->P.R || C > b->P.C) return PP_INVALID_ARGS;
	if (reserve == 0) reserve = std::max<uint64_t>(4ull * n0, 1);
	if (reserve < n0 || reserve > b->P.cap) return PP_INVALID_ARGS;
	if (!b->ctl_fresh) PP_TRY(read_ctl(b));
	if (b->ctl_host->error) return b->ctl_host->error;
	const uint32_t tick = b->ctl_host->tick, pos = b->ctl_host->n;
	/* (a plain simulation runs without a limit; pp_sim_run checks its runs) */
	if (iters != ITERS_UNLIMITED &&
		(uint64_t)tick + iters + b->P.R + 1 >= 0xFFFFFFFFull)
		return PP_INVALID_ARGS;
	if (b->reserved + reserve > b->P.cap || (uint64_t)pos + n0 > b->P.cap)
		return PP_OUT_OF_RESOURCES;
	int slot = -1;
	for (uint32_t s = 0; s < b->max_rep; s++)
		if (b->slot_rid[s] < 0) {
			slot = (int)s;
			break;
		}
	if (slot < 0) return PP_OUT_OF_RESOURCES;
	const int64_t cb = take_cells(b, C);
	if (cb < 0) return PP_OUT_OF_RESOURCES;

	Rep rp{};
	rp.cb = (uint32_t)cb;
	rp.C = C;
	rp.gx = p->grid_x;
	div_magic(rp.gx, &rp.gx_magic, &rp.gx_shift);
	rp.gy = p->grid_y;
	rp.R = p->grass_restart;
	rp.sheep_gain = p->sheep_gain_from_food;
	rp.wolves_gain = p->wolves_gain_from_food;
	rp.sheep_thr = p->sheep_reproduce_threshold;
	rp.wolves_thr = p->wolves_reproduce_threshold;
	rp.sheep_prob = p->sheep_reproduce_prob;
	rp.wolves_prob = p->wolves_reproduce_prob;
	rp.init_sheep = p->init_sheep;
	rp.n0 = n0;
	rp.key0 = (uint32_t)seed;
	rp.key1 = (uint32_t)(seed >> 32);
	rp.t0 = tick;
	rp.iters = iters;
	rp.cap = (uint32_t)reserve;
	rp.stats_rows = b->stats_rows;
	rp.ring = b->ring + (size_t)slot * b->ring_stride;
	rp.stats = b->stats_dev + (size_t)slot * b->stats_rows;
	int rc = PP_SUCCESS;
	if (hipMemcpy(&b->rep_dev[slot], &rp, sizeof(Rep), hipMemcpyHostToDevice)
			!= hipSuccess)
		rc = PP_LIBRARY_ERROR;
	if (rc) {
		release_cells(b, (uint32_t)cb, C);
		return rc;
	}
	b->rep_host[slot] = rp;
	const uint32_t id = b->next_id++;
	b->slot_rid[slot] = id;
	b->reserved += reserve;
	Replica &r = b->reps[id];
	r.id = id;
	r.slot = slot;
	r.cb = (uint32_t)cb;
	r.C = C;
	r.prm = *p;
	r.prm.grid_xy = C;
	r.seed = seed;
	r.iters = iters;
	r.reserve = (uint32_t)reserve;
	r.n0 = n0;
	r.status = RS_ACTIVE;
	r.error = 0;
	r.err_step = 0xFFFFFFFFu;
	r.rows_have = 0;
	r.n = n0;
	r.agent_steps = 0;
	r.keep = keep;
	r.sink = nullptr;
	r.sink_from = 0;
	PP_TRY(upload_table(b));

	const Tables T = b->tables();
	const uint32_t cu = (uint32_t)b->num_cu;
	PendingEvent pe;
	prof_begin(b, KI_INIT_RESET, pe);
	k_rep_reset<<<1, WG, 0, b->stream>>>(T, (uint32_t)slot);
	prof_end(b, pe);
	prof_begin(b, KI_INIT_CELLS, pe);
	const uint32_t gc = min_u32(ceil_div(C, WG), 8 * cu);
	if (b->grass_bytes == 1)
		k_init_cells<uint8_t><<<gc, WG, 0, b->stream>>>((uint8_t *)b->grass_at, T,
			(uint32_t)slot);
	else if (b->grass_bytes == 2)
		k_init_cells<uint16_t><<<gc, WG, 0, b->stream>>>((uint16_t *)b->grass_at,
			T, (uint32_t)slot);
	else
		k_init_cells<uint32_t><<<gc, WG, 0, b->stream>>>((uint32_t *)b->grass_at,
			T, (uint32_t)slot);
	prof_end(b, pe);
	prof_begin(b, KI_INIT_AGENTS, pe);
	k_init_agents<<<min_u32(ceil_div(n0 ? n0 : 1, WG), 8 * cu), WG, 0,
		b->stream>>>(b->cell, b->energy, b->type, T, (uint32_t)slot, pos);
	prof_end(b, pe);
	prof_begin(b, KI_INIT_DONE, pe);
	k_rep_init_done<<<1, 64, 0, b->stream>>>(b->ctl, T, (uint32_t)slot, pos);
	prof_end(b, pe);
	prof_begin(b, KI_MOVE0, pe);
	k_move0<<<min_u32(ceil_div(n0 ? n0 : 1, WG), 8 * cu), WG, 0,
		b->stream>>>(b->cell, b->energy, b->type, b->keys[0], b->vals[0],
		b->ctl, T, (uint32_t)slot, pos, b->P);
	prof_end(b, pe);
	HIP_OK(hipGetLastError());
	b->ctl_fresh = 0;
	if (id_out) *id_out = id;
	return PP_SUCCESS;
}

/* Batch positions of a replica's current agents: (base, n); the stream must
 * be idle and the mirrors fresh. */
static int state_of(pp_batch_t *b, const Replica &r, uint32_t *cell,
	int32_t *energy, uint8_t *type, uint32_t *grass) {
	if (r.slot < 0) return PP_INVALID_ARGS;
	const RepCtl &rc = b->rctl_host[r.slot];
	const size_t base = rc.base, n = rc.n;
	if (cell && n) {
		HIP_OK(hipMemcpy(cell, b->cell + base, n * 4, hipMemcpyDeviceToHost));
		for (size_t i = 0; i < n; i++) cell[i] -= r.cb;
	}
	if (energy && n)
		HIP_OK(hipMemcpy(energy, b->energy + base, n * 4, hipMemcpyDeviceToHost));
	if (type && n) HIP_OK(hipMemcpy(type, b->type + base, n, hipMemcpyDeviceToHost));
	if (grass) {
		/* Countdown at tick t from the regrow tick a stored modulo 2^w:
		 * d = (a - t) mod 2^w is the countdown if 1 <= d <= R, else 0. */
		const uint32_t t = b->ctl_host->tick, R = r.prm.grass_restart, C = r.C;
		const size_t gb = b->grass_bytes;
		std::vector<uint8_t> raw((size_t)C * gb);
		HIP_OK(hipMemcpy(raw.data(), (const uint8_t *)b->grass_at + (size_t)r.cb * gb,
			raw.size(), hipMemcpyDeviceToHost));
		const uint32_t mask = gb == 4 ? 0xFFFFFFFFu : ((1u << (8 * gb)) - 1u);
		for (uint32_t c = 0; c < C; c++) {
			uint32_t a;
			if (gb == 1) a = raw[c];
			else if (gb == 2) a = reinterpret_cast<const uint16_t *>(raw.data())[c];
			else a = reinterpret_cast<const uint32_t *>(raw.data())[c];
			const uint32_t d = (a - t) & mask;
			grass[c] = (d >= 1 && d <= R) ? d : 0;
		}
	}
	return PP_SUCCESS;
}

static int prof_get(pp_batch_t *b, int i, const char **name,
	uint64_t *launches, double *total_ns) {
	if (i < 0 || i >= KI_COUNT) return PP_INVALID_ARGS;
	PP_TRY(prof_collect(b));
	if (name) *name = kernel_names[i];
	if (launches) *launches = b->prof_launches[i];
	if (total_ns) *total_ns = b->prof_ns[i];
	return PP_SUCCESS;
}

static int prof_export(pp_batch_t *b, const char *path) {
	PP_TRY(prof_collect(b));
	FILE *fp = fopen(path, "w");
	if (!fp) return PP_UNABLE_SAVE_STATS;
	double total = 0;
	for (int i = 0; i < KI_COUNT; i++) total += b->prof_ns[i];
	for (int i = 0; i < KI_COUNT; i++)
		fprintf(fp, "\"%s\"\t%llu\t%lg\n", kernel_names[i],
			(unsigned long long)b->prof_ns[i],
			total > 0 ? b->prof_ns[i] / total : 0.0);
	fclose(fp);
	return PP_SUCCESS;
}

static double now_s() {
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Times `ticks` ticks of the current shape (graph replay, then collect). */
static int time_ticks(pp_batch_t *b, uint32_t ticks, double *sec) {
	HIP_OK(hipStreamSynchronize(b->stream));
	const double t0 = now_s();
	PP_TRY(enqueue_ticks(b, ticks));
	HIP_OK(hipStreamSynchronize(b->stream));
	*sec = now_s() - t0;
	return collect(b);
}

/* ------------------------------------------------ allocation-free runs */

/* Limits of pp_sim_simulate_noalloc: one plain simulation of at most
 * NOALLOC_MAX_AGENTS agents (its capacity) on at most NOALLOC_MAX_CELLS cells
 * with grass restart at most NOALLOC_MAX_R, every device buffer a slice of
 * `scratch` (about 10 MiB at the limits). */
constexpr uint64_t NOALLOC_MAX_AGENTS = 1u << 18;
constexpr uint64_t NOALLOC_MAX_CELLS = 1u << 18;
constexpr uint32_t NOALLOC_MAX_R = 4096;
constexpr size_t NOALLOC_BYTES = 12u << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

/* The device scratch of pp_sim_simulate_noalloc. Each GPU holds its own copy
 * until the process ends. In host code the name is a host-side stand-in, not
 * a device address: host code uses only the address hipGetSymbolAddress
 * returns for the current device (noalloc_scratch). */
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static void *g_scratch_dev[NOALLOC_MAX_DEVICES];   /* its address per device */
static std::mutex g_scratch_mu;
/* Host copies of the control words of the run on each device. */
static Ctl g_na_ctl[NOALLOC_MAX_DEVICES];
static RepCtl g_na_rctl[NOALLOC_MAX_DEVICES];

/* scratch's device address on device dev (current), looked up on the first
 * call on that device and kept. */
static int noalloc_scratch(int dev, unsigned char **out) {
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_dev[dev]) {
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) {
			fprintf(stderr, "pphpc: hipGetSymbolAddress(scratch) on device %d: "
				"%s\n", dev, hipGetErrorName(e));
			return PP_LIBRARY_ERROR;
		}
		g_scratch_dev[dev] = p;
	}
	*out = (unsigned char *)g_scratch_dev[dev];
	return PP_SUCCESS;
}

/* pp_sim_create + pp_sim_simulate + pp_sim_destroy with every buffer carved
 * from the scratch region, on the null stream, without graphs: no device or
 * pinned-host allocation and no heap allocation on the host (the batch record
 * lives on the stack and its containers stay empty). Rows 0..iters go to
 * rows. */
static int simulate_noalloc(const pp_params_t *p, uint64_t seed,
	uint64_t capacity, int device, pp_stats64_t *rows) {
	if (!p || !rows) return PP_INVALID_ARGS;
	PP_TRY(check_params(p));
	const uint64_t C = (uint64_t)p->grid_x * p->grid_y;
	const uint64_t n0 = (uint64_t)p->init_sheep + p->init_wolves;
	if (capacity == 0) capacity = std::max<uint64_t>(4 * n0, 1);
	if (capacity < n0) return PP_INVALID_ARGS;
	if (capacity > NOALLOC_MAX_AGENTS || C > NOALLOC_MAX_CELLS ||
		p->grass_restart > NOALLOC_MAX_R) {
		fprintf(stderr, "pphpc: pp_sim_simulate_noalloc: capacity %llu, %llu "
			"cells, grass restart %u; the limits are %llu agents, %llu cells, "
			"grass restart %u\n", (unsigned long long)capacity,
			(unsigned long long)C, p->grass_restart,
			(unsigned long long)NOALLOC_MAX_AGENTS,
			(unsigned long long)NOALLOC_MAX_CELLS, NOALLOC_MAX_R);
		return PP_INVALID_ARGS;
	}
	if ((uint64_t)p->iters + p->grass_restart + 1 >= 0xFFFFFFFFull)
		return PP_INVALID_ARGS;
	int dev = device;
	if (dev < 0) HIP_OK(hipGetDevice(&dev));
	if (dev >= NOALLOC_MAX_DEVICES) return PP_INVALID_ARGS;
	HIP_OK(hipSetDevice(dev));
	unsigned char *base = nullptr;
	PP_TRY(noalloc_scratch(dev, &base));

	pp_batch_t bb;
	pp_batch_t *b = &bb;
	b->device = dev;
	if (hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount,
			dev) != hipSuccess || b->num_cu <= 0)
		b->num_cu = 256;
	b->stream = nullptr;
	b->no_graph = 1;
	b->P.cap = (uint32_t)capacity;
	b->P.C = (uint32_t)C;
	b->P.R = p->grass_restart;
	b->P.nb_cap = nb_cap_for(C);
	b->P.bshift = 11;
	b->sort_mode = 1;
	b->max_rep = 1;
	b->stats_rows = STATS_RING;
	b->key_bits = key_bits_for(C);
	b->grass_bytes = grass_bytes_for(p->grass_restart);
	apply_shape(b, default_shape(b->key_bits));
	b->hist_alloc_words = hist_bytes_for(b->shape, b->key_bits, b->P.cap) / 4;
	size_t off = 0;
	PP_TRY(batch_buffers(b, [&](void **ptr, size_t bytes) {
		off = (off + 255) & ~(size_t)255;
		*ptr = off + bytes <= NOALLOC_BYTES ? base + off : base;
		off += bytes;
		return hipSuccess;
	}));
	if (off > NOALLOC_BYTES) {
		/* (the limits above keep every layout inside) */
		fprintf(stderr, "pphpc: pp_sim_simulate_noalloc: %zu bytes of scratch "
			"needed, %zu available\n", off, NOALLOC_BYTES);
		return PP_INVALID_ARGS;
	}
	b->ctl_host = &g_na_ctl[dev];
	b->rctl_host = &g_na_rctl[dev];

	/* The scratch keeps the last call's contents: clear what must start at
	 * zero (control block, partial sums, histograms, the replica's control
	 * words, the bucket counts). */
	const hipStream_t st = b->stream;
	k_batch_reset<<<min_u32(ceil_div(b->hist_alloc_words, WG), 2048), WG, 0,
		st>>>(b->ctl, b->hist, (uint32_t)b->hist_alloc_words, b->P.part);
	HIP_OK(hipGetLastError());
	HIP_OK(hipMemsetAsync(b->rctl_dev, 0, sizeof(RepCtl), st));
	const size_t nbw = cnt_words(b->P.nb_cap);
	HIP_OK(hipMemsetAsync(b->ocnt, 0, nbw * 4, st));
	HIP_OK(hipMemsetAsync(b->ncnt, 0, nbw * 4, st));

	/* The replica in slot 0, cells from 0 (as admit and upload_table do
	 * for the one replica of a plain simulation). */
	Rep rp{};
	rp.cb = 0;
	rp.C = (uint32_t)C;
	rp.gx = p->grid_x;
	div_magic(rp.gx, &rp.gx_magic, &rp.gx_shift);
	rp.gy = p->grid_y;
	rp.R = p->grass_restart;
	rp.sheep_gain = p->sheep_gain_from_food;
	rp.wolves_gain = p->wolves_gain_from_food;
	rp.sheep_thr = p->sheep_reproduce_threshold;
	rp.wolves_thr = p->wolves_reproduce_threshold;
	rp.sheep_prob = p->sheep_reproduce_prob;
	rp.wolves_prob = p->wolves_reproduce_prob;
	rp.init_sheep = p->init_sheep;
	rp.n0 = (uint32_t)n0;
	rp.key0 = (uint32_t)seed;
	rp.key1 = (uint32_t)(seed >> 32);
	rp.t0 = 0;
	rp.iters = p->iters;
	rp.cap = (uint32_t)capacity;
	rp.stats_rows = b->stats_rows;
	rp.ring = b->ring;
	rp.stats = b->stats_dev;
	const uint32_t shift = 11;
	const uint32_t h = rows_per_bucket(rp.gx, shift);
	const uint64_t nbk = (uint64_t)((rp.gy + h - 1) / h) * row_blocks(rp.gx, shift);
	const uint32_t on = (nbk <= b->P.nb_cap && nbk <= CNT_MAX_BUCKETS) ? 1u : 0u;
	uint32_t top = 0;
	while (top + 1 < (uint32_t)CNT_LEVELS && (nbk >> (6 * (top + 1))) > 0) top++;
	b->P.bucket_on = on;
	b->P.bshift = shift;
	b->P.cnt_top = top;
	b->P.one_slot = 0;
	b->P.solo = 1;
	const uint32_t zero = 0, one = 1, nb32 = on ? (uint32_t)nbk : 0u;
	const uint32_t bbase[2] = {0, (uint32_t)nbk};
	HIP_OK(hipMemcpy(b->rep_dev, &rp, sizeof(Rep), hipMemcpyHostToDevice));
	HIP_OK(hipMemcpy(b->cbs_dev, &zero, 4, hipMemcpyHostToDevice));
	HIP_OK(hipMemcpy(b->order_dev, &zero, 4, hipMemcpyHostToDevice));
	HIP_OK(hipMemcpy(b->bbase_dev, bbase, 8, hipMemcpyHostToDevice));
	HIP_OK(hipMemcpy(&b->ctl->nrep, &one, 4, hipMemcpyHostToDevice));
	HIP_OK(hipMemcpy(&b->ctl->nbuckets, &nb32, 4, hipMemcpyHostToDevice));
	b->need_radix = 1;

	const Tables T = b->tables();
	const uint32_t cu = (uint32_t)b->num_cu;
	k_rep_reset<<<1, WG, 0, st>>>(T, 0);
	const uint32_t gc = min_u32(ceil_div(C, WG), 8 * cu);
	if (b->grass_bytes == 1)
		k_init_cells<uint8_t><<<gc, WG, 0, st>>>((uint8_t *)b->grass_at, T, 0);
	else if (b->grass_bytes == 2)
		k_init_cells<uint16_t><<<gc, WG, 0, st>>>((uint16_t *)b->grass_at, T, 0);
	else
		k_init_cells<uint32_t><<<gc, WG, 0, st>>>((uint32_t *)b->grass_at, T, 0);
	const uint32_t ga = min_u32(ceil_div(n0 ? n0 : 1, WG), 8 * cu);
	k_init_agents<<<ga, WG, 0, st>>>(b->cell, b->energy, b->type, T, 0, 0);
	k_rep_init_done<<<1, 64, 0, st>>>(b->ctl, T, 0, 0);
	k_move0<<<ga, WG, 0, st>>>(b->cell, b->energy, b->type, b->keys[0],
		b->vals[0], b->ctl, T, 0, 0, b->P);
	HIP_OK(hipGetLastError());
	HIP_OK(hipMemcpy(&rows[0], b->stats_dev, sizeof(pp_stats64_t),
		hipMemcpyDeviceToHost));

	/* Chunks of at most half the stats ring, rows copied after each. */
	uint32_t done = 0;
	while (done < p->iters) {
		const uint32_t chunk = min_u32(p->iters - done, b->stats_rows / 2);
		PP_TRY(enqueue_ticks(b, chunk));
		HIP_OK(hipMemcpy(b->ctl_host, b->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
		HIP_OK(hipMemcpy(b->rctl_host, b->rctl_dev, sizeof(RepCtl),
			hipMemcpyDeviceToHost));
		if (b->ctl_host->error) return b->ctl_host->error;
		const RepCtl &rc = *b->rctl_host;
		const uint32_t last = rc.step;   /* rows 0..last exist */
		for (uint32_t i = done + 1; i <= last; i++)
			HIP_OK(hipMemcpy(&rows[i], b->stats_dev + i % b->stats_rows,
				sizeof(pp_stats64_t), hipMemcpyDeviceToHost));
		if (rc.status == RS_FAILED) return rc.error;
		if (last != done + chunk) return PP_LIBRARY_ERROR;
		done = last;
	}
	return PP_SUCCESS;
}

/* Runs the body of an entry point: no C++ exception crosses the C ABI (an
 * allocation that fails becomes PP_OUT_OF_RESOURCES, anything else
 * PP_LIBRARY_ERROR). */
template <typename F>
static int guarded(F &&f) {
	try {
		return f();
	} catch (const std::bad_alloc &) {
		return PP_OUT_OF_RESOURCES;
	} catch (...) {
		return PP_LIBRARY_ERROR;
	}
}

extern "C" {

/* ------------------------------------------------------ single simulation */

int pp_sim_create(const pp_params_t *p, uint64_t seed, uint64_t capacity,
	int device, pp_sim_t **out) {
	return guarded([&]() -> int {
		if (!p || !out) return PP_INVALID_ARGS;
		*out = nullptr;
		PP_TRY(check_params(p));
		const uint64_t C = (uint64_t)p->grid_x * p->grid_y;
		const uint64_t n0 = (uint64_t)p->init_sheep + p->init_wolves;
		if (capacity == 0) {
			capacity = 4 * n0;
			if (capacity < PP_DEFAULT_MAX_AGENTS) capacity = PP_DEFAULT_MAX_AGENTS;
		}
		if (capacity < n0 || capacity >= (1ull << 29)) return PP_INVALID_ARGS;
		pp_sim_t *s = new pp_sim_t();
		int rc = batch_create(capacity, C, 1, p->grass_restart, device, &s->b);
		if (rc) {
			delete s;
			return rc;
		}
		s->prm = *p;
		s->prm.grid_xy = (uint32_t)C;
		s->seed = seed;
		s->cap = (uint32_t)capacity;
		*out = s;
		return PP_SUCCESS;
	});
}

void pp_sim_destroy(pp_sim_t *s) {
	if (!s) return;
	batch_free(s->b);
	delete s;
}

static Replica *sim_rep(const pp_sim_t *s) {
	if (!s || !s->initialised) return nullptr;
	auto it = s->b->reps.find(s->rid);
	return it == s->b->reps.end() ? nullptr : &it->second;
}

int pp_sim_init(pp_sim_t *s, pp_stats64_t *stats0) {
	return guarded([&]() -> int {
		if (!s) return PP_INVALID_ARGS;
		s->initialised = 0;
		PP_TRY(batch_reset(s->b));
		uint32_t id = 0;
		PP_TRY(admit(s->b, &s->prm, s->seed, s->cap, ITERS_UNLIMITED, 0, &id));
		s->rid = id;
		s->initialised = 1;
		Replica *r = sim_rep(s);
		r->sink = stats0;
		r->sink_from = 0;
		const int rc = collect(s->b);
		r->sink = nullptr;
		if (rc) return rc;
		if (s->b->ctl_host->error) return s->b->ctl_host->error;
		return PP_SUCCESS;
	});
}

int pp_sim_run(pp_sim_t *s, uint32_t iters, pp_stats64_t *out) {
	return guarded([&]() -> int {
		Replica *r = sim_rep(s);
		if (!r) return PP_INVALID_ARGS;
		pp_batch_t *b = s->b;
		HIP_OK(hipSetDevice(b->device));
		if (b->ctl_host->error) return b->ctl_host->error;
		if (r->error) return r->error;
		const uint64_t last = (uint64_t)b->ctl_host->tick + iters + b->P.R + 1;
		if (last >= 0xFFFFFFFFull) return PP_INVALID_ARGS;
		uint32_t done = 0;
		while (done < iters) {
			const uint32_t chunk = min_u32(iters - done, b->stats_rows / 2);
			const uint32_t step0 = r->rows_have - 1;
			r->sink = out ? out + done : nullptr;
			r->sink_from = step0 + 1;
			int rc = enqueue_ticks(b, chunk);
			if (!rc) rc = collect(b);
			r->sink = nullptr;
			if (rc) return rc;
			const uint32_t got = r->rows_have - 1 - step0;
			done += got;
			if (b->ctl_host->error) return b->ctl_host->error;
			if (r->error) return r->error;
			if (got != chunk) return PP_LIBRARY_ERROR;
		}
		return PP_SUCCESS;
	});
}

int pp_sim_step(pp_sim_t *s, pp_stats64_t *out) {
	return pp_sim_run(s, 1, out);
}

int pp_sim_simulate(pp_sim_t *s, pp_stats64_t *out) {
	return guarded([&]() -> int {
		if (!s) return PP_INVALID_ARGS;
		PP_TRY(pp_sim_init(s, out));
		return pp_sim_run(s, s->prm.iters, out ? out + 1 : nullptr);
	});
}

uint32_t pp_sim_step_count(const pp_sim_t *s) {
	const Replica *r = sim_rep(s);
	return r ? r->rows_have - 1 : 0;
}

uint64_t pp_sim_num_agents(const pp_sim_t *s) {
	const Replica *r = sim_rep(s);
	return r ? r->n : 0;
}

uint32_t pp_sim_error_step(const pp_sim_t *s) {
	if (!s) return 0xFFFFFFFFu;
	if (s->b->ctl_host->error) return s->b->ctl_host->err_step;
	const Replica *r = sim_rep(s);
	return (r && r->error) ? r->err_step : 0xFFFFFFFFu;
}

uint64_t pp_sim_agent_steps(const pp_sim_t *s) {
	return s ? s->b->ctl_host->agent_steps : 0;
}

uint64_t pp_sim_capacity(const pp_sim_t *s) { return s ? s->cap : 0; }

int pp_sim_counters(const pp_sim_t *s, uint64_t out[6]) {
	if (!s || !out) return PP_INVALID_ARGS;
	const Ctl *c = s->b->ctl_host;
	out[0] = c->agent_steps;
	out[1] = c->sum_live;
	out[2] = c->sum_out;
	out[3] = c->sum_occ;
	out[4] = c->sum_eaten;
	out[5] = c->sum_births;
	return PP_SUCCESS;
}

int pp_sim_get_state(pp_sim_t *s, uint32_t *cell, int32_t *energy,
	uint8_t *type, uint32_t *grass) {
	return guarded([&]() -> int {
		Replica *r = sim_rep(s);
		if (!r) return PP_INVALID_ARGS;
		HIP_OK(hipSetDevice(s->b->device));
		PP_TRY(read_ctl(s->b));
		return state_of(s->b, *r, cell, energy, type, grass);
	});
}

int pp_sim_dump(pp_sim_t *s, const char *agents_path, const char *cells_path,
	uint32_t dump_type, int append) {
	return guarded([&]() -> int {
		Replica *r = sim_rep(s);
		if (!r) return PP_INVALID_ARGS;
		HIP_OK(hipSetDevice(s->b->device));
		PP_TRY(read_ctl(s->b));
		if (r->slot < 0) return PP_INVALID_ARGS;
		const uint64_t n = s->b->rctl_host[r->slot].n;
		const uint64_t cells = (uint64_t)s->prm.grid_x * s->prm.grid_y;
		std::vector<uint32_t> cell(n), grass(cells);
		std::vector<int32_t> energy(n);
		std::vector<uint8_t> type(n);
		PP_TRY(state_of(s->b, *r, cell.data(), energy.data(), type.data(),
			grass.data()));
		return pp_dump_state(agents_path, cells_path, append, r->rows_have - 1,
			dump_type, &s->prm, n, cell.data(), energy.data(), type.data(),
			grass.data());
	});
}

int pp_sim_set_profiling(pp_sim_t *s, int enable) {
	if (!s) return PP_INVALID_ARGS;
	s->b->profiling = enable ? 1 : 0;
	return PP_SUCCESS;
}

int pp_sim_prof_count(const pp_sim_t *s) { return s ? KI_COUNT : 0; }

int pp_sim_prof_get(pp_sim_t *s, int i, const char **name, uint64_t *launches,
	double *total_ns) {
	return guarded([&]() -> int {
		if (!s) return PP_INVALID_ARGS;
		return prof_get(s->b, i, name, launches, total_ns);
	});
}

int pp_sim_prof_reset(pp_sim_t *s) {
	return guarded([&]() -> int {
		if (!s) return PP_INVALID_ARGS;
		return pp_batch_prof_reset(s->b);
	});
}

int pp_sim_prof_export(pp_sim_t *s, const char *path) {
	return guarded([&]() -> int {
		if (!s || !path) return PP_INVALID_ARGS;
		return prof_export(s->b, path);
	});
}

int pp_sim_info(const pp_sim_t *s, uint32_t *sort_passes, uint32_t *sort_tile,
	uint32_t *action_tile, uint32_t *grid_move) {
	if (!s) return PP_INVALID_ARGS;
	return pp_batch_info(s->b, sort_passes, sort_tile, action_tile, grid_move);
}

int pp_sim_set_sort(pp_sim_t *s, uint32_t bucket_shift) {
	return guarded([&]() -> int {
		if (!s) return PP_INVALID_ARGS;
		return pp_batch_set_sort(s->b, bucket_shift);
	});
}

int pp_sim_sync(pp_sim_t *s) {
	if (!s) return PP_INVALID_ARGS;
	return pp_batch_sync(s->b);
}

/* ---------------------------------------------------------------- batches */

int pp_sim_simulate_noalloc(const pp_params_t *p, uint64_t seed,
	uint64_t capacity, int device, pp_stats64_t *rows) {
	return guarded([&]() -> int {
		return simulate_noalloc(p, seed, capacity, device, rows);
	});
}

int pp_batch_create(uint64_t agent_capacity, uint64_t cell_capacity,
	uint32_t max_replicas, uint32_t max_grass_restart, int device,
	pp_batch_t **out) {
	return guarded([&]() -> int {
		if (!out) return PP_INVALID_ARGS;
		return batch_create(agent_capacity, cell_capacity, max_replicas,
			max_grass_restart, device, out);
	});
}

void pp_batch_destroy(pp_batch_t *b) { batch_free(b); }

uint64_t pp_batch_footprint(uint64_t agent_capacity, uint64_t cell_capacity,
	uint32_t max_replicas, uint32_t max_grass_restart) {
	try {
		if (agent_capacity == 0 || cell_capacity == 0 || max_replicas == 0 ||
			max_grass_restart == 0 || agent_capacity >= (1ull << 29) ||
			cell_capacity > 0xFFFFFFFEull)
			return 0;
		return batch_bytes(agent_capacity, cell_capacity, max_replicas,
			max_grass_restart, stats_rows_for(max_replicas),
			max_hist_words(key_bits_for(cell_capacity), (uint32_t)agent_capacity));
	} catch (...) {
		return 0;
	}
}

int pp_batch_add(pp_batch_t *b, const pp_params_t *p, uint64_t seed,
	uint64_t agent_reserve, uint32_t *id) {
	return guarded([&]() -> int {
		if (!b || !p) return PP_INVALID_ARGS;
		/* (UINT32_MAX is the unlimited count of a plain simulation) */
		if (p->iters == ITERS_UNLIMITED) return PP_INVALID_ARGS;
		return admit(b, p, seed, agent_reserve, p->iters, 1, id);
	});
}

int pp_batch_run(pp_batch_t *b, uint32_t ticks) {
	return guarded([&]() -> int {
		if (!b) return PP_INVALID_ARGS;
		return advance(b, ticks);
	});
}

int pp_batch_replica(pp_batch_t *b, uint32_t id, int *state, uint32_t *steps,
	int *error, uint64_t *agent_steps) {
	if (!b) return PP_INVALID_ARGS;
	auto it = b->reps.find(id);
	if (it == b->reps.end()) return PP_INVALID_ARGS;
	const Replica &r = it->second;
	if (state) *state = (int)r.status;
	if (steps) *steps = r.rows_have ? r.rows_have - 1 : 0;
	if (error) *error = r.error;
	if (agent_steps) *agent_steps = r.agent_steps;
	return PP_SUCCESS;
}

int pp_batch_rows(pp_batch_t *b, uint32_t id, uint32_t first, uint32_t n,
	pp_stats64_t *out) {
	if (!b || (n && !out)) return PP_INVALID_ARGS;
	auto it = b->reps.find(id);
	if (it == b->reps.end()) return PP_INVALID_ARGS;
	const Replica &r = it->second;
	if ((uint64_t)first + n > r.rows_have) return PP_INVALID_ARGS;
	for (uint32_t i = 0; i < n; i++) out[i] = r.rows[first + i];
	return PP_SUCCESS;
}

int pp_batch_get_state(pp_batch_t *b, uint32_t id, uint32_t *n_agents,
	uint32_t *cell, int32_t *energy, uint8_t *type, uint32_t *grass) {
	return guarded([&]() -> int {
		if (!b) return PP_INVALID_ARGS;
		auto it = b->reps.find(id);
		if (it == b->reps.end()) return PP_INVALID_ARGS;
		HIP_OK(hipSetDevice(b->device));
		PP_TRY(read_ctl(b));
		const Replica &r = it->second;
		if (r.slot < 0) return PP_INVALID_ARGS;
		if (n_agents) *n_agents = b->rctl_host[r.slot].n;
		return state_of(b, r, cell, energy, type, grass);
	});
}

int pp_batch_release(pp_batch_t *b, uint32_t id) {
	if (!b) return PP_INVALID_ARGS;
	auto it = b->reps.find(id);
	if (it == b->reps.end() || it->second.slot >= 0) return PP_INVALID_ARGS;
	b->reps.erase(it);
	return PP_SUCCESS;
}

int pp_batch_resources(const pp_batch_t *b, uint64_t *agents_free,
	uint32_t *cells_free_run, uint32_t *slots_free, uint32_t *active) {
	if (!b) return PP_INVALID_ARGS;
	if (agents_free) *agents_free = b->P.cap - b->reserved;
	if (cells_free_run) {
		uint32_t m = 0;
		for (auto &kv : b->free_cells) m = std::max(m, kv.second);
		*cells_free_run = m;
	}
	if (slots_free) {
		uint32_t n = 0;
		for (int64_t v : b->slot_rid) n += v < 0;
		*slots_free = n;
	}
	if (active) *active = active_count(b);
	return PP_SUCCESS;
}

uint32_t pp_batch_tick(const pp_batch_t *b) {
	return b ? b->ctl_host->tick : 0;
}

int pp_batch_counters(const pp_batch_t *b, uint64_t out[6]) {
	if (!b || !out) return PP_INVALID_ARGS;
	const Ctl *c = b->ctl_host;
	out[0] = c->agent_steps;
	out[1] = c->sum_live;
	out[2] = c->sum_out;
	out[3] = c->sum_occ;
	out[4] = c->sum_eaten;
	out[5] = c->sum_births;
	return PP_SUCCESS;
}

int pp_batch_info(const pp_batch_t *b, uint32_t *sort_passes,
	uint32_t *sort_tile, uint32_t *action_tile, uint32_t *grid_move) {
	if (!b) return PP_INVALID_ARGS;
	if (sort_passes) *sort_passes = b->P.passes;
	if (sort_tile) *sort_tile = b->P.sort_tile;
	if (action_tile) *action_tile = A_TILE;
	if (grid_move) *grid_move = b->grid_emit;
	return PP_SUCCESS;
}

int pp_batch_set_shape(pp_batch_t *b, uint32_t keys_per_thread,
	uint32_t max_digit_bits) {
	return guarded([&]() -> int {
		if (!b) return PP_INVALID_ARGS;
		if (keys_per_thread != 8 && keys_per_thread != 16) return PP_INVALID_ARGS;
		if (max_digit_bits != 8 && max_digit_bits != 11) return PP_INVALID_ARGS;
		const SortShape sh{(int)keys_per_thread, (int)max_digit_bits};
		if (hist_bytes_for(sh, b->key_bits, b->P.cap) > b->hist_alloc_words * 4)
			return PP_INVALID_ARGS;
		HIP_OK(hipSetDevice(b->device));
		if (sh.kpt == b->shape.kpt && sh.max_bits == b->shape.max_bits)
			return PP_SUCCESS;
		apply_shape(b, sh);
		return PP_SUCCESS;
	});
}

int pp_batch_set_launch(pp_batch_t *b, uint32_t tile_rounds,
	uint32_t bucket_eighths) {
	return guarded([&]() -> int {
		if (!b) return PP_INVALID_ARGS;
		if (tile_rounds != 1 && tile_rounds != 2 && tile_rounds != 4 &&
			tile_rounds != 8)
			return PP_INVALID_ARGS;
		if (bucket_eighths < 1 || bucket_eighths > 8) return PP_INVALID_ARGS;
		HIP_OK(hipSetDevice(b->device));
		if (tile_rounds != b->tile_rounds || bucket_eighths != b->bucket_eighths)
			apply_launch(b, tile_rounds, bucket_eighths);
		return PP_SUCCESS;
	});
}

int pp_batch_get_launch(const pp_batch_t *b, uint32_t *tile_rounds,
	uint32_t *bucket_eighths) {
	if (!b) return PP_INVALID_ARGS;
	if (tile_rounds) *tile_rounds = b->tile_rounds;
	if (bucket_eighths) *bucket_eighths = b->bucket_eighths;
	return PP_SUCCESS;
}

int pp_batch_set_sort(pp_batch_t *b, uint32_t mode) {
	return guarded([&]() -> int {
		if (!b) return PP_INVALID_ARGS;
		if (mode > 1 && (mode < 8 || mode > 11)) return PP_INVALID_ARGS;
		HIP_OK(hipSetDevice(b->device));
		if (mode == b->sort_mode) return PP_SUCCESS;
		if (!b->ctl_fresh) PP_TRY(read_ctl(b));
		b->sort_mode = mode;
		return upload_table(b);    /* new layout; the next tick re-sorts */
	});
}

int pp_batch_tune(pp_batch_t *b, uint32_t trial_ticks, uint32_t *bucket_shift,
	uint32_t *keys_per_thread, uint32_t *max_digit_bits, double *sec_per_tick) {
	return guarded([&]() -> int {
		if (!b || trial_ticks == 0 || trial_ticks + 8 > b->stats_rows / 2)
			return PP_INVALID_ARGS;
		HIP_OK(hipSetDevice(b->device));
		/* candidates: the bucket sort at each block width, then the radix sort
		 * at each shape */
		struct Cand { uint32_t shift; SortShape sh; };
		std::vector<Cand> cands;
		for (uint32_t sft = 8; sft <= 11; sft++) cands.push_back({sft, b->shape});
		for (const SortShape &sh : tune_candidates(b->key_bits, b->P.cap))
			cands.push_back({0, sh});
		const uint32_t mode0 = b->sort_mode;
		const SortShape shape0 = b->shape;
		Cand best{b->P.bucket_on ? b->P.bshift : 0u, b->shape};
		double best_t = -1.0;
		/* One trial of the current shapes: 8 untimed ticks re-sort if
		 * needed, build the graph, warm caches; a trial during which
		 * replicas finished (the table changed) is settled and timed once
		 * more. */
		auto trial = [&](double &t) -> int {
			for (int attempt = 0; attempt < 2; attempt++) {
				PP_TRY(enqueue_ticks(b, 8));
				PP_TRY(collect(b));
				const uint64_t v0 = b->table_version;
				PP_TRY(time_ticks(b, trial_ticks, &t));
				if (b->table_version == v0 || active_count(b) == 0) break;
			}
			return PP_SUCCESS;
		};
		for (const Cand &c : cands) {
			if (active_count(b) == 0) break;
			PP_TRY(pp_batch_set_sort(b, c.shift));
			PP_TRY(pp_batch_set_shape(b, (uint32_t)c.sh.kpt, (uint32_t)c.sh.max_bits));
			if (c.shift && !b->P.bucket_on) continue;   /* layout too fine */
			double t = 0;
			PP_TRY(trial(t));
			if (best_t < 0 || t < best_t) {
				best_t = t;
				best = c;
			}
		}
		if (best_t < 0) {
			/* nothing was timed: the sort as it was */
			PP_TRY(pp_batch_set_sort(b, mode0));
			PP_TRY(pp_batch_set_shape(b, (uint32_t)shape0.kpt,
				(uint32_t)shape0.max_bits));
		} else {
			/* the automatic choice when it is the winner (it then keeps
			 * adapting to later replica sets), else the winner pinned */
			const SortShape def = default_shape(b->key_bits);
			PP_TRY(pp_batch_set_sort(b, 1));
			PP_TRY(pp_batch_set_shape(b, (uint32_t)def.kpt, (uint32_t)def.max_bits));
			const bool auto_wins = b->P.bucket_on ? best.shift == 11
				: (best.shift == 0 && best.sh.kpt == def.kpt &&
					best.sh.max_bits == def.max_bits);
			if (!auto_wins) {
				PP_TRY(pp_batch_set_sort(b, best.shift));
				PP_TRY(pp_batch_set_shape(b, (uint32_t)best.sh.kpt,
					(uint32_t)best.sh.max_bits));
			}
		}
		/* Then the launch shape under the chosen sort: rounds of resident
		 * groups of the action and emit grids, and the bucket sort's half
		 * or whole round. (The emit width is left to its rule, which follows
		 * the replica set: choosing it here, in trials with every replica
		 * running, made the 64-replica workload's run 8% slower.) */
		if (best_t >= 0) {
			uint32_t best_r = b->tile_rounds, best_e = b->bucket_eighths;
			double best_l = -1.0;
			for (uint32_t r : {1u, 2u, 4u, 8u})
				for (uint32_t e8 : {4u, 8u}) {
					if (active_count(b) == 0) break;
					if (!b->P.bucket_on && e8 != BUCKET_EIGHTHS) continue;
					apply_launch(b, r, e8);
					double t = 0;
					PP_TRY(trial(t));
					if (best_l < 0 || t < best_l) {
						best_l = t;
						best_r = r;
						best_e = e8;
					}
				}
			apply_launch(b, best_r, best_e);
			if (best_l >= 0) best_t = best_l;
		}
		if (bucket_shift) *bucket_shift = best.shift;
		if (keys_per_thread) *keys_per_thread = (uint32_t)best.sh.kpt;
		if (max_digit_bits) *max_digit_bits = (uint32_t)best.sh.max_bits;
		if (sec_per_tick) *sec_per_tick = best_t < 0 ? 0.0 : best_t / trial_ticks;
		return b->ctl_host->error;
	});
}

int pp_batch_set_profiling(pp_batch_t *b, int enable) {
	if (!b) return PP_INVALID_ARGS;
	b->profiling = enable ? 1 : 0;
	return PP_SUCCESS;
}

int pp_batch_prof_count(const pp_batch_t *b) { return b ? KI_COUNT : 0; }

int pp_batch_prof_get(pp_batch_t *b, int i, const char **name,
	uint64_t *launches, double *total_ns) {
	return guarded([&]() -> int {
		if (!b) return PP_INVALID_ARGS;
		return prof_get(b, i, name, launches, total_ns);
	});
}

int pp_batch_prof_reset(pp_batch_t *b) {
	return guarded([&]() -> int {
		if (!b) return PP_INVALID_ARGS;
		const int rc = prof_collect(b);
		memset(b->prof_launches, 0, sizeof(b->prof_launches));
		memset(b->prof_ns, 0, sizeof(b->prof_ns));
		return rc;
	});
}

int pp_batch_sync(pp_batch_t *b) {
	if (!b) return PP_INVALID_ARGS;
	HIP_OK(hipSetDevice(b->device));
	HIP_OK(hipStreamSynchronize(b->stream));
	return PP_SUCCESS;
}

} /* extern "C" */
