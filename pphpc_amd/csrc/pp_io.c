/*
 * pp_io.c -- parameter-file parser and statistics writer of the PPHPC step.
 *
 * Keeps the reference's file formats: the 12-key "KEY = value" parameter file
 * read by pp_load_params (opencl/pp/pp_common.c:39-195) and the tab-separated
 * per-iteration statistics written by pp_stats_save (pp_common.c:205-240);
 * and the text state dump of pp_dump_state, after the reference's PPG_DUMP
 * output (pp_gpu.c:438-555).
 */
#include "pphpc.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Keys in the order of the PPParameters fields (pp_common.in.h:129-157),
 * grid_xy excluded; their bit in the "seen" mask is their index, as in the
 * reference's check word (pp_common.c:56-169). */
static const char *const pp_keys[12] = {
	"INIT_SHEEP",
	"SHEEP_GAIN_FROM_FOOD",
	"SHEEP_REPRODUCE_THRESHOLD",
	"SHEEP_REPRODUCE_PROB",
	"INIT_WOLVES",
	"WOLVES_GAIN_FROM_FOOD",
	"WOLVES_REPRODUCE_THRESHOLD",
	"WOLVES_REPRODUCE_PROB",
	"GRASS_RESTART",
	"GRID_X",
	"GRID_Y",
	"ITERS",
};

int pp_load_params(const char *path, pp_params_t *out) {
	uint32_t field[12] = {0};
	uint32_t seen = 0;
	char key[100];
	int value;
	FILE *fp;

	if (!out) return PP_INVALID_ARGS;
	fp = fopen(path ? path : "config.txt", "r");
	if (!fp) return PP_UNABLE_TO_OPEN_PARAMS_FILE;

	/* One "%s = %d" record per scan, as pp_common.c:54 reads them. */
	while (fscanf(fp, "%99s = %d", key, &value) != EOF) {
		int idx = -1;
		for (int k = 0; k < 12; k++) {
			if (strcmp(key, pp_keys[k]) == 0) { idx = k; break; }
		}
		if (idx < 0 || (seen >> idx) & 1u) {
			/* Unknown or repeated key (pp_common.c:60-62, :163-166). */
			fclose(fp);
			return PP_INVALID_PARAMS_FILE;
		}
		seen |= 1u << idx;
		field[idx] = (uint32_t)value;
	}
	fclose(fp);
	if (seen != 0xfffu) return PP_INVALID_PARAMS_FILE; /* pp_common.c:169 */

	out->init_sheep = field[0];
	out->sheep_gain_from_food = field[1];
	out->sheep_reproduce_threshold = field[2];
	out->sheep_reproduce_prob = field[3];
	out->init_wolves = field[4];
	out->wolves_gain_from_food = field[5];
	out->wolves_reproduce_threshold = field[6];
	out->wolves_reproduce_prob = field[7];
	out->grass_restart = field[8];
	out->grid_x = field[9];
	out->grid_y = field[10];
	out->grid_xy = field[9] * field[10]; /* pp_common.c:177 */
	out->iters = field[11];
	return PP_SUCCESS;
}

int pp_stats_save(const char *path, const pp_stats64_t *rows, uint32_t nrows,
	const pp_params_t *p) {
	FILE *fp;
	if (!p || (nrows && !rows)) return PP_INVALID_ARGS;
	fp = fopen(path ? path : "stats.txt", "w");
	if (!fp) return PP_UNABLE_SAVE_STATS;
	for (uint32_t i = 0; i < nrows; i++) {
		const pp_stats64_t *r = rows + i;
		double sheep_mean = r->sheep ? (double)r->sheep_en / (double)r->sheep : 0.0;
		double wolf_mean = r->wolves ? (double)r->wolves_en / (double)r->wolves : 0.0;
		double grass_mean = (double)r->grass_en / (double)p->grid_xy;
		if (fprintf(fp, "%llu\t%llu\t%llu\t%f\t%f\t%f\n",
				(unsigned long long)r->sheep, (unsigned long long)r->wolves,
				(unsigned long long)r->grass, sheep_mean, wolf_mean,
				grass_mean) < 0) {
			fclose(fp);
			return PP_UNABLE_SAVE_STATS;
		}
	}
	if (fclose(fp) != 0) return PP_UNABLE_SAVE_STATS;
	return PP_SUCCESS;
}

int pp_stats_save_f32(const char *path, const pp_stats64_t *rows,
	uint32_t nrows, const pp_params_t *p) {
	FILE *fp;
	if (!p || (nrows && !rows)) return PP_INVALID_ARGS;
	fp = fopen(path ? path : "stats.txt", "w");
	if (!fp) return PP_UNABLE_SAVE_STATS;
	for (uint32_t i = 0; i < nrows; i++) {
		const pp_stats64_t *r = rows + i;
		/* the reference's cl_uint fields (pp_common.in.h:109-123) */
		const uint32_t sheep = (uint32_t)r->sheep, wolves = (uint32_t)r->wolves;
		const uint32_t grass = (uint32_t)r->grass, sheep_en = (uint32_t)r->sheep_en;
		const uint32_t wolves_en = (uint32_t)r->wolves_en;
		const uint32_t grass_en = (uint32_t)r->grass_en;
		/* single-precision quotients, as pp_common.c:218-224 computes them */
		const float sheep_mean = sheep > 0 ? (float)sheep_en / (float)sheep : 0.0f;
		const float wolf_mean = wolves > 0 ? (float)wolves_en / (float)wolves : 0.0f;
		const float grass_mean = (float)grass_en / (float)p->grid_xy;
		if (fprintf(fp, "%d\t%d\t%d\t%f\t%f\t%f\n", (int)sheep, (int)wolves,
				(int)grass, (double)sheep_mean, (double)wolf_mean,
				(double)grass_mean) < 0) {
			fclose(fp);
			return PP_UNABLE_SAVE_STATS;
		}
	}
	if (fclose(fp) != 0) return PP_UNABLE_SAVE_STATS;
	return PP_SUCCESS;
}

const char *pp_strerror(int code) {
	switch (code) {
	case PP_SUCCESS: return "success";
	case PP_UNKNOWN_ARGS: return "unknown arguments";
	case PP_INVALID_ARGS: return "invalid arguments";
	case PP_LIBRARY_ERROR: return "error in HIP runtime";
	case PP_UNABLE_TO_OPEN_PARAMS_FILE: return "unable to open parameters file";
	case PP_INVALID_PARAMS_FILE: return "invalid parameters file";
	case PP_UNABLE_SAVE_STATS: return "unable to save statistics";
	case PP_OUT_OF_RESOURCES: return "agents above capacity (out of resources)";
	case PP_ENERGY_OVERFLOW: return "agent energy above its 24-bit limit";
	default: return "unknown error code";
	}
}

int pp_dump_state(const char *agents_path, const char *cells_path, int append,
	uint32_t iter, uint32_t dump_type, const pp_params_t *p, uint64_t n,
	const uint32_t *cell, const int32_t *energy, const uint8_t *type,
	const uint32_t *grass) {
	if (!p || !p->grid_x || (n && (!cell || !energy || !type)) ||
		(cells_path && !grass))
		return PP_INVALID_ARGS;
	const uint64_t cells = (uint64_t)p->grid_x * p->grid_y;
	const char *mode = append ? "a" : "w";
	if (agents_path) {
		FILE *fp = fopen(agents_path, mode);
		if (!fp) return PP_UNABLE_SAVE_STATS;
		fprintf(fp, "\nIter %u, agents=%llu\n", iter, (unsigned long long)n);
		for (uint64_t k = 0; k < n; k++)
			fprintf(fp, "[%4llu] (%4u, %4u) type=%u energy=%d\n",
				(unsigned long long)k, cell[k] % p->grid_x, cell[k] / p->grid_x,
				(unsigned)type[k], energy[k]);
		if (fclose(fp)) return PP_UNABLE_SAVE_STATS;
	}
	if (cells_path) {
		/* first and last index of each cell's agents */
		uint64_t *first = malloc(cells * sizeof(uint64_t));
		uint64_t *last = malloc(cells * sizeof(uint64_t));
		if (!first || !last) {
			free(first);
			free(last);
			return PP_OUT_OF_RESOURCES;
		}
		for (uint64_t c = 0; c < cells; c++) first[c] = UINT64_MAX;
		for (uint64_t k = 0; k < n; k++) {
			if (cell[k] >= cells) {
				free(first);
				free(last);
				return PP_INVALID_ARGS;
			}
			if (first[cell[k]] == UINT64_MAX) first[cell[k]] = k;
			last[cell[k]] = k;
		}
		FILE *fp = fopen(cells_path, mode);
		if (!fp) {
			free(first);
			free(last);
			return PP_UNABLE_SAVE_STATS;
		}
		fprintf(fp, "\nIteration %u\n", iter);
		int blank_line = 0;
		for (uint64_t c = 0; c < cells; c++) {
			const unsigned x = (unsigned)(c % p->grid_x);
			const unsigned y = (unsigned)(c / p->grid_x);
			if ((dump_type & 0x10) && first[c] == UINT64_MAX) {
				blank_line = 1;
				continue;
			}
			if (blank_line) fprintf(fp, "\n");
			blank_line = 0;
			if (first[c] == UINT64_MAX)
				fprintf(fp, "(%u, %u) -> (-, -)  [Grass: %u]\n", x, y, grass[c]);
			else
				fprintf(fp, "(%u, %u) -> (%llu, %llu) %s [Grass: %u]\n", x, y,
					(unsigned long long)first[c], (unsigned long long)last[c],
					first[c] != last[c] ? "More than 1 agent present" : "",
					grass[c]);
		}
		free(first);
		free(last);
		if (fclose(fp)) return PP_UNABLE_SAVE_STATS;
	}
	return PP_SUCCESS;
}
