/*
 * pp_main.c -- the `pp` executable: drop-in for the reference GPU simulator's
 * command line (opencl/pp/pp_gpu.c:310-345, main at :1700-1948).
 *
 *   pp [-p FILE] [-s FILE] [-r SEED] [-m MAX_AGENTS] [-d INDEX] [-i FILE]
 *      [-D PREFIX [-T TYPE]]
 *
 * -p/--params     parameters file (default config.txt)
 * -s/--stats      statistics output file (default stats.txt)
 * -r/--rng-seed   RNG seed (default 0)
 * -m/--max-agents agent capacity (default max(4 x initial agents, 16777216))
 * -d/--device     HIP device index (default: current device)
 * -i/--prof-agg   export per-kernel times as "name"\tns\trelative
 *                 (pp_common.c:248-260)
 * -D/--dump       write PREFIX_agents.txt and PREFIX_cells.txt: the state
 *                 after init and after every iteration, as the reference's
 *                 PPG_DUMP build writes dump_agents.txt / dump_cells.txt
 *                 (pp_gpu.c:36-46, :644-657, :924-931); runs the iterations
 *                 one at a time
 * -T/--dump-type  PPG_DUMP's type bits (default 0x01; 0x10: only the cells
 *                 holding agents)
 *
 * The number of iterations is ITERS of the parameters file. The exit status
 * is the PP_* code of the first error (pp_common.in.h:87-104), 0 on success.
 */
#define _POSIX_C_SOURCE 200809L
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pphpc.h"

static double now_s(void) {
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void usage(FILE *fp) {
	fprintf(fp,
		"Usage: pp [-p FILE] [-s FILE] [-r SEED] [-m MAX_AGENTS] [-d INDEX] "
		"[-i FILE] [-D PREFIX [-T TYPE]] [-F]\n"
		"  -p, --params FILE      parameters file (default config.txt)\n"
		"  -s, --stats FILE       statistics output (default stats.txt)\n"
		"  -r, --rng-seed SEED    random seed (default 0)\n"
		"  -m, --max-agents N     agent capacity\n"
		"  -d, --device INDEX     HIP device index\n"
		"  -i, --prof-agg FILE    export aggregate per-kernel times\n"
		"  -D, --dump PREFIX      dump the state of every iteration to\n"
		"                         PREFIX_agents.txt and PREFIX_cells.txt\n"
		"  -T, --dump-type TYPE   dump type bits (0x10: cells with agents)\n"
		"  -F, --f32-stats        statistics as the reference prints them\n"
		"                         (32-bit counts, single-precision means)\n");
}

/* pp_sim_simulate one iteration at a time, dumping every state. */
static int simulate_dumping(pp_sim_t *sim, const pp_params_t *p,
	pp_stats64_t *rows, const char *prefix, uint32_t dump_type) {
	char *agents = malloc(strlen(prefix) + 16), *cells = malloc(strlen(prefix) + 16);
	if (!agents || !cells) {
		free(agents);
		free(cells);
		return PP_OUT_OF_RESOURCES;
	}
	sprintf(agents, "%s_agents.txt", prefix);
	sprintf(cells, "%s_cells.txt", prefix);
	int rc = pp_sim_init(sim, &rows[0]);
	if (!rc) rc = pp_sim_dump(sim, agents, cells, dump_type, 0);
	for (uint32_t i = 1; !rc && i <= p->iters; i++) {
		rc = pp_sim_step(sim, &rows[i]);
		if (!rc) rc = pp_sim_dump(sim, agents, cells, dump_type, 1);
	}
	free(agents);
	free(cells);
	return rc;
}

int main(int argc, char **argv) {
	static const struct option opts[] = {
		{"params", required_argument, NULL, 'p'},
		{"stats", required_argument, NULL, 's'},
		{"rng-seed", required_argument, NULL, 'r'},
		{"max-agents", required_argument, NULL, 'm'},
		{"device", required_argument, NULL, 'd'},
		{"prof-agg", required_argument, NULL, 'i'},
		{"dump", required_argument, NULL, 'D'},
		{"dump-type", required_argument, NULL, 'T'},
		{"f32-stats", no_argument, NULL, 'F'},
		{"help", no_argument, NULL, 'h'},
		{NULL, 0, NULL, 0}};
	const char *params_file = NULL, *stats_file = NULL, *prof_file = NULL;
	const char *dump_prefix = NULL;
	unsigned long dump_type = 0x01;
	unsigned long long seed = 0, cap = 0;
	int device = -1, opt, f32_stats = 0;
	char *endp;

	while ((opt = getopt_long(argc, argv, "p:s:r:m:d:i:D:T:Fh", opts, NULL)) != -1) {
		switch (opt) {
		case 'p': params_file = optarg; break;
		case 's': stats_file = optarg; break;
		case 'r':
			seed = strtoull(optarg, &endp, 0);
			if (*endp) { fprintf(stderr, "Error: invalid seed\n"); return PP_INVALID_ARGS; }
			break;
		case 'm':
			cap = strtoull(optarg, &endp, 0);
			if (*endp) { fprintf(stderr, "Error: invalid -m\n"); return PP_INVALID_ARGS; }
			break;
		case 'd':
			device = (int)strtol(optarg, &endp, 0);
			if (*endp) { fprintf(stderr, "Error: invalid -d\n"); return PP_INVALID_ARGS; }
			break;
		case 'i': prof_file = optarg; break;
		case 'D': dump_prefix = optarg; break;
		case 'T':
			dump_type = strtoul(optarg, &endp, 0);
			if (*endp) { fprintf(stderr, "Error: invalid -T\n"); return PP_INVALID_ARGS; }
			break;
		case 'F': f32_stats = 1; break;
		case 'h': usage(stdout); return PP_SUCCESS;
		default: usage(stderr); return PP_UNKNOWN_ARGS;
		}
	}
	if (optind < argc) {
		/* The reference refuses non-option arguments (pp_common.c:295-307). */
		fprintf(stderr, "Error: this program does not accept non-option arguments.\n");
		return PP_UNKNOWN_ARGS;
	}

	pp_params_t p;
	int rc = pp_load_params(params_file, &p);
	if (rc) {
		fprintf(stderr, "Error: %s (\"%s\")\n", pp_strerror(rc),
			params_file ? params_file : "config.txt");
		return rc;
	}

	pp_sim_t *sim = NULL;
	rc = pp_sim_create(&p, seed, cap, device, &sim);
	if (rc) {
		fprintf(stderr, "Error: %s\n", pp_strerror(rc));
		return rc;
	}
	if (prof_file) pp_sim_set_profiling(sim, 1);

	pp_stats64_t *rows = calloc((size_t)p.iters + 1, sizeof(pp_stats64_t));
	if (!rows) {
		pp_sim_destroy(sim);
		return PP_OUT_OF_RESOURCES;
	}
	uint32_t passes = 0, tile = 0, atile = 0, gmove = 0;
	pp_sim_info(sim, &passes, &tile, &atile, &gmove);
	printf("PPHPC step on HIP: %ux%u cells, %u+%u agents, capacity %llu, "
		"%u iterations, %u radix passes (tile %u)\n",
		p.grid_x, p.grid_y, p.init_sheep, p.init_wolves,
		(unsigned long long)pp_sim_capacity(sim), p.iters, passes, tile);

	double t0 = now_s();
	if (dump_prefix)
		rc = simulate_dumping(sim, &p, rows, dump_prefix, (uint32_t)dump_type);
	else
		rc = pp_sim_simulate(sim, rows);
	double t1 = now_s();
	if (rc) {
		if (rc == PP_OUT_OF_RESOURCES)
			fprintf(stderr, "Error: agents above capacity at iteration %u "
				"(capacity %llu)\n", pp_sim_error_step(sim),
				(unsigned long long)pp_sim_capacity(sim));
		else if (rc == PP_ENERGY_OVERFLOW)
			fprintf(stderr, "Error: agent energy above %u at iteration %u\n",
				PP_ENERGY_MAX, pp_sim_error_step(sim));
		else
			fprintf(stderr, "Error: %s\n", pp_strerror(rc));
	} else {
		rc = f32_stats ? pp_stats_save_f32(stats_file, rows, p.iters + 1, &p)
			: pp_stats_save(stats_file, rows, p.iters + 1, &p);
		if (rc) fprintf(stderr, "Error: %s\n", pp_strerror(rc));
		if (prof_file && !rc) {
			rc = pp_sim_prof_export(sim, prof_file);
			if (rc) fprintf(stderr, "Error: %s\n", pp_strerror(rc));
		}
		printf("Elapsed time: %.4es (%.4e agent-steps/s)\n", t1 - t0,
			(double)pp_sim_agent_steps(sim) / (t1 - t0));
	}
	free(rows);
	pp_sim_destroy(sim);
	return rc;
}
