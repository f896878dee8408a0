/*
 * pp_philox.h -- Philox4x32-10 counter-based RNG for gfx950 device code.
 *
 * Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
 * (SC'11): 10 rounds of the Philox-4x32 S-box with the Weyl key schedule.
 * Replaces the per-work-item stateful generators the reference takes from
 * cl_ops (clo_rng_next_int, call sites opencl/pp/pp_gpu.cl:383-433, :760,
 * :814, :951): every draw is a pure function of (seed, counter), so results do
 * not depend on thread mapping, launch shape or execution order.
 *
 * Counter layout used by the step: (agent id or cell, step, stream, block).
 */
#ifndef PP_PHILOX_H
#define PP_PHILOX_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PP_STREAM_INIT_CELL 0u
#define PP_STREAM_INIT_AGENT 1u
#define PP_STREAM_MOVE 2u
#define PP_STREAM_SHUFFLE 3u

__device__ __forceinline__ uint4 pp_philox(uint32_t c0, uint32_t c1,
	uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
	for (int r = 0; r < 10; r++) {
		/* one 32 x 32 -> 64 product per multiplier (a single
		 * v_mad_u64_u32 instead of separate high and low multiplies) */
		const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
		const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
		const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
		const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
		const uint32_t n0 = hi1 ^ c1 ^ k0;
		const uint32_t n2 = hi0 ^ c3 ^ k1;
		c0 = n0;
		c1 = lo1;
		c2 = n2;
		c3 = lo0;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	return make_uint4(c0, c1, c2, c3);
}

/* Uniform integer in [0, n): floor(x * n / 2^32), integer only. */
__device__ __forceinline__ uint32_t pp_bounded(uint32_t x, uint32_t n) {
	return __umulhi(x, n);
}

#endif
