/*
 * sampo_abms.h -- the reference driver's run (src/abms.c:464-555 with the
 * set-up of main, :586-705) as one call over the icl boundary of
 * libsampo_icl.so.  It keeps the reference's kernel sequence and launch
 * widths: it is the plain, one-kernel-per-step GPU path.
 */
#ifndef SAMPO_ABMS_H
#define SAMPO_ABMS_H

#include <stdint.h>
#include "sampo_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	uint32_t seed;          /* srand() seed (src/abms.c:611) */
	uint32_t n_eggs;        /* initialAgentCount (src/abms.c:18) */
	uint32_t max_agents;    /* maxAgents (include/abms.h:16) */
	uint32_t steps;         /* maxSteps (src/abms.c:16) */
	uint32_t serial_scan;   /* 1: CPU-device scan path of src/abms.c:416-439 */
	const float *temp;      /* steps temperatures (readSingleValFile, src/abms.c:94-108) */
	const sampo_env *env;   /* steps environments or 1 (readEnvironment, src/abms.c:110-137) */
	uint32_t env_count;
	/* configuration mode (SAMPO_MODE_CONFIG): the given initial population in
	 * place of initAgents' eggs, and constant per-stage daily mortality dmr
	 * (egg, larva, pupa, adult) in place of killAgents' formulas */
	uint32_t mode;
	float dmr[4];
	const sampo_layout *init_layout;
	const sampo_agent *init_agents; /* dense in range order */
	/* the speciesHeader argument (src/abms.c:565-568): NULL or a name ending in
	 * properties.h for include/properties.h, else a species file path */
	const char *species;
} sampo_abms_params;

typedef struct {
	sampo_row *rows;        /* steps rows (printStats, src/abms.c:67-81), may be NULL */
	uint32_t status;        /* 0, SAMPO_CAPACITY or SAMPO_EXTINCT */
	uint32_t steps_done;
	uint64_t agent_steps;   /* agents alive at the start of each iteration, summed */
	sampo_layout layout;    /* final popD[1] ranges (printPopulation, src/abms.c:83-92) */
	uint32_t histogram[100];/* storeHistogram (src/abms.c:160-183) */
	double loop_ms;         /* wall time of the iteration loop */
	double total_ms;        /* wall time of the whole call */
	/* optional copies of the final agent buffers (24 / 16 / 8 bytes per slot) */
	void *agents, *ages, *states;
	uint64_t span_capacity;
} sampo_abms_result;

int sampo_abms_run(const sampo_abms_params *p, sampo_abms_result *res);

#ifdef __cplusplus
}
#endif
#endif
