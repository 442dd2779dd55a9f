/*
 * sampo_abms.c -- the reference driver's iteration loop (src/abms.c) on the
 * icl boundary of libsampo_icl.so.
 *
 * Same functions, same order, same kernel launches and launch widths as
 * src/abms.c:140-555; differences: the seed of the host seed stream is a
 * parameter (src/abms.c:611 seeds from the clock), temperature and
 * environment arrays come from the caller instead of files, and the stats
 * rows, final population, histogram and agent arrays are returned instead of
 * printed.  DEVICE_TYPE GPU path: preScan + scan (src/abms.c:404-415).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sampo_icl.h"
#include "../../include/sampo_abms.h"
#include "sampo_canon.h"
#include "sampo_model.h"

#define LOCAL_SIZE 64u
#define NSEEDS 4u

typedef struct { cl_uint start, end; } Range;
typedef struct {
	Range r[8];
	cl_uint numLarvae1DayEquiv, numFemale, numPotentiallyInfective;
} Population;
typedef struct { cl_uint z1, z2, z3, z4; } Seeds;
typedef struct { cl_uint numCycles, sumCycles, numBites, numInfectBites; } Bnc;
typedef struct { cl_int totalBiomass; cl_uint newEggs; } Enb;

enum { EGG = 1, LARVA = 2, PUPA = 4, IMMATURE = 8, MATESEEKING = 16, BMS = 32, BMD = 64, GRAVID = 128 };

typedef struct {
	icl_device *dev;
	const sampo_abms_params *p;
	sc_glibc_rand rng;
	cl_uint maxAgents;
	icl_kernel *calcL1dePerGroup, *calcL1deTotal, *calcFemalesPerGroup, *calcFemalesTotal;
	icl_kernel *killAgents, *updateAgents, *oldToNewAgents, *preScan, *serialScan, *createEggs;
	float hmr[4]; /* configuration mode: hourly mortality egg, larva, pupa, adult */
	char opts[1100]; /* kernelBuildArgs (src/abms.c:565-568) */
} Driver;

static size_t roundup(size_t v, size_t m) { return ((v + m - 1) / m) * m; }

/* src/abms.c:140-158 */
static void generate_seeds(Driver *d, icl_buffer *seedsD, Seeds *seedsH)
{
	for (cl_uint i = 0; i < NSEEDS; ++i) {
		seedsH[i].z1 = sc_glibc_next(&d->rng);
		seedsH[i].z2 = sc_glibc_next(&d->rng);
		seedsH[i].z3 = sc_glibc_next(&d->rng);
		seedsH[i].z4 = sc_glibc_next(&d->rng);
	}
	icl_write_buffer(seedsD, CL_TRUE, NSEEDS * sizeof(Seeds), seedsH, NULL, NULL);
}

/* Configuration mode: the given agents written into the reference's
 * array-of-structs buffers at their padded indices (agent.h:66-102); gap
 * slots are zero and never valid. */
static void upload_population(Driver *d, icl_buffer *agents, icl_buffer *ages, icl_buffer *states, Population *pop)
{
	const sampo_layout *L = d->p->init_layout;
	const cl_uint span = L->end[7];
	typedef struct { cl_uint availableEggs, cycleLength, humanBloodmealCount; float delay; cl_uint numEggBatches, aux; } A;
	typedef struct { float ageInHours, hoursInState; cl_uint isFemale; float csd; } G;
	typedef struct { cl_uint dead, state; } St;
	A *ha = (A *)calloc(span ? span : 1, sizeof(A));
	G *hg = (G *)calloc(span ? span : 1, sizeof(G));
	St *hs = (St *)calloc(span ? span : 1, sizeof(St));
	ICL_ASSERT(ha && hg && hs, "host allocation failed");
	uint64_t k = 0;
	for (int s = 0; s < 8; ++s) {
		pop->r[s].start = L->start[s];
		pop->r[s].end = L->end[s];
		for (cl_uint g = L->start[s]; g < L->end[s]; ++g, ++k) {
			const sampo_agent *a = &d->p->init_agents[k];
			ha[g].availableEggs = a->availableEggs;
			ha[g].cycleLength = a->cycleLength;
			ha[g].humanBloodmealCount = a->humanBloodmealCount;
			ha[g].delay = a->delay;
			ha[g].numEggBatches = a->numEggBatches;
			ha[g].aux = a->aux;
			hg[g].ageInHours = a->ageInHours;
			hg[g].hoursInState = a->hoursInState;
			hg[g].isFemale = a->isFemale;
			hg[g].csd = a->csd;
			hs[g].dead = 0u;
			hs[g].state = 1u << s;
		}
	}
	if (span) {
		icl_write_buffer(agents, CL_TRUE, (size_t)span * sizeof(A), ha, NULL, NULL);
		icl_write_buffer(ages, CL_TRUE, (size_t)span * sizeof(G), hg, NULL, NULL);
		icl_write_buffer(states, CL_TRUE, (size_t)span * sizeof(St), hs, NULL, NULL);
	}
	free(ha);
	free(hg);
	free(hs);
}

/* src/abms.c:185-228 */
static void create_initial_population(Driver *d, icl_buffer *agents, icl_buffer *ages, icl_buffer *states,
                                      Population *pop, icl_buffer *enbD, icl_buffer *seeds)
{
	d->createEggs = icl_create_kernel(d->dev, "kernel/initAgents.cl", "initAgents", d->opts, ICL_SOURCE);
	if (d->p->mode == SAMPO_MODE_CONFIG) {
		upload_population(d, agents, ages, states, pop);
		return;
	}
	size_t local = LOCAL_SIZE;
	size_t global = roundup(d->maxAgents, LOCAL_SIZE);
	float t0 = d->p->temp[0];
	icl_run_kernel(d->createEggs, 1, &global, &local, NULL, NULL, 6, (size_t)0, (void *)agents, (size_t)0,
	               (void *)ages, (size_t)0, (void *)states, (size_t)0, (void *)enbD, sizeof(float), &t0,
	               (size_t)0, (void *)seeds);
	pop->r[0].start = 0;
	pop->r[0].end = d->p->n_eggs;
	for (int s = 1; s < 8; ++s) pop->r[s].start = pop->r[s].end = d->p->n_eggs + 1u;
}

/* src/abms.c:230-310; returns 0, or SAMPO_CAPACITY / SAMPO_EXTINCT */
static int calc_stats(Driver *d, icl_buffer *ages, icl_buffer *popD, icl_buffer *buff, Population *popH,
                      sampo_row *row)
{
	size_t local = LOCAL_SIZE;
	size_t global = LOCAL_SIZE * LOCAL_SIZE;
	cl_uint slots = LOCAL_SIZE;
	icl_run_kernel(d->calcL1dePerGroup, 1, &global, &local, NULL, NULL, 3, (size_t)0, (void *)ages, (size_t)0,
	               (void *)popD, (size_t)0, (void *)buff);
	icl_run_kernel(d->calcL1deTotal, 1, &local, &local, NULL, NULL, 3, (size_t)0, (void *)popD, (size_t)0,
	               (void *)buff, sizeof(cl_uint), &slots);
	icl_run_kernel(d->calcFemalesPerGroup, 1, &global, &local, NULL, NULL, 3, (size_t)0, (void *)ages, (size_t)0,
	               (void *)popD, (size_t)0, (void *)buff);
	icl_run_kernel(d->calcFemalesTotal, 1, &local, &local, NULL, NULL, 2, (size_t)0, (void *)popD, (size_t)0,
	               (void *)buff);
	icl_read_buffer(popD, CL_TRUE, sizeof(Population), popH, NULL, NULL);
	cl_uint n = 0;
	for (int s = 0; s < 8; ++s) {
		row->count[s] = popH->r[s].end - popH->r[s].start;
		n += row->count[s];
	}
	cl_uint upper = popH->r[7].end;
	cl_uint adults = upper - popH->r[3].start;
	row->l1de = row->count[1] > 0 ? popH->numLarvae1DayEquiv : 0u;
	row->females = adults > 0 ? popH->numFemale : 0u;
	row->potInfective = adults > 0 ? popH->numPotentiallyInfective : 0u;
	row->valid = 1u;
	if (!((double)n <= (double)d->maxAgents * 0.95)) return SAMPO_CAPACITY;
	if (!(n > 0)) return SAMPO_EXTINCT;
	return 0;
}

/* src/abms.c:312-341 */
static void update(Driver *d, icl_buffer *agents, icl_buffer *ages, icl_buffer *states, icl_buffer *enbD,
                   icl_buffer *bnc, icl_buffer *popD, icl_buffer *seeds, cl_uint worldTime, const Population *popH,
                   cl_uint i)
{
	size_t local = LOCAL_SIZE;
	size_t global = roundup(popH->r[7].end, LOCAL_SIZE);
	const sampo_env *env = &d->p->env[d->p->env_count == 1 ? 0 : i];
	float cc = env->carryingCapacity;
	float t = d->p->temp[i];
	float dt = 1.0f;
	if (d->p->mode == SAMPO_MODE_CONFIG)
		icl_run_kernel(d->killAgents, 1, &global, &local, NULL, NULL, 10, (size_t)0, (void *)ages, (size_t)0,
		               (void *)states, (size_t)0, (void *)popD, (size_t)0, (void *)enbD, (size_t)0, (void *)bnc,
		               (size_t)0, (void *)seeds, sizeof(float), &d->hmr[0], sizeof(float), &d->hmr[1],
		               sizeof(float), &d->hmr[2], sizeof(float), &d->hmr[3]);
	else
		icl_run_kernel(d->killAgents, 1, &global, &local, NULL, NULL, 7, (size_t)0, (void *)ages, (size_t)0,
		               (void *)states, (size_t)0, (void *)popD, (size_t)0, (void *)enbD, (size_t)0, (void *)bnc,
		               (size_t)0, (void *)seeds, sizeof(float), &cc);
	icl_run_kernel(d->updateAgents, 1, &global, &local, NULL, NULL, 11, (size_t)0, (void *)agents, (size_t)0,
	               (void *)ages, (size_t)0, (void *)states, (size_t)0, (void *)popD, sizeof(sampo_env), (void *)env,
	               sizeof(float), &t, (size_t)0, (void *)bnc, (size_t)0, (void *)enbD, (size_t)0, (void *)seeds,
	               sizeof(float), &dt, sizeof(cl_uint), &worldTime);
}

/* src/abms.c:358-379 */
static void parallel_prefix_scan(Driver *d, icl_buffer *states, icl_buffer *pop, icl_buffer *prefix,
                                 icl_buffer *flags, cl_uint start, cl_uint b1, cl_uint b2, cl_uint end, cl_int match)
{
	size_t local = LOCAL_SIZE;
	size_t global = ((end + LOCAL_SIZE) / LOCAL_SIZE) * LOCAL_SIZE;
	icl_run_kernel(d->preScan, 1, &global, &local, NULL, NULL, 9, (size_t)0, (void *)states, (size_t)0, (void *)pop,
	               (size_t)0, (void *)prefix, (size_t)0, (void *)flags, sizeof(cl_uint), &start, sizeof(cl_uint), &b1,
	               sizeof(cl_uint), &b2, sizeof(cl_uint), &end, sizeof(cl_int), &match);
	scan(prefix, flags, end + 1 - start);
}

/* src/abms.c:343-356 */
static void sequential_prefix_scan(Driver *d, icl_buffer *states, icl_buffer *pop, icl_buffer *prefix,
                                   cl_uint offset, cl_uint start, cl_uint end, cl_uint match)
{
	size_t one = 1;
	icl_run_kernel(d->serialScan, 1, &one, &one, NULL, NULL, 7, (size_t)0, (void *)states, (size_t)0,
	               (void *)prefix, (size_t)0, (void *)pop, sizeof(cl_uint), &offset, sizeof(cl_uint), &start,
	               sizeof(cl_uint), &end, sizeof(cl_uint), &match);
}

/* src/abms.c:381-462 */
static void create_new_agents(Driver *d, icl_buffer *agents, icl_buffer *ages, icl_buffer *states,
                              icl_buffer *nagents, icl_buffer *nages, icl_buffer *nstates, icl_buffer *enbD,
                              icl_buffer *popD, icl_buffer *seeds, icl_buffer *p1, icl_buffer *p2, icl_buffer *p3,
                              icl_buffer *buff, const Population *popH, cl_uint i)
{
	size_t local = LOCAL_SIZE;
	size_t global = ((d->maxAgents + LOCAL_SIZE) / LOCAL_SIZE) * LOCAL_SIZE;
	float t = d->p->temp[i];
	icl_run_kernel(d->createEggs, 1, &global, &local, NULL, NULL, 6, (size_t)0, (void *)nagents, (size_t)0,
	               (void *)nages, (size_t)0, (void *)nstates, (size_t)0, (void *)enbD, sizeof(float), &t, (size_t)0,
	               (void *)seeds);
	if (!d->p->serial_scan) {
		parallel_prefix_scan(d, states, popD, p1, buff, 0, popH->r[0].end, popH->r[3].end, popH->r[6].end,
		                     EGG | IMMATURE | BMD);
		parallel_prefix_scan(d, states, popD, p2, buff, 0, popH->r[1].end, popH->r[4].end, popH->r[7].end,
		                     LARVA | MATESEEKING | GRAVID);
		parallel_prefix_scan(d, states, popD, p3, buff, popH->r[1].start, popH->r[2].end, popH->r[7].end,
		                     popH->r[7].end, PUPA | BMS);
	} else {
		cl_uint ls = popH->r[1].start;
		sequential_prefix_scan(d, states, popD, p1, 0, popH->r[0].start, popH->r[0].end, EGG);
		sequential_prefix_scan(d, states, popD, p2, 0, popH->r[0].start, popH->r[1].end, LARVA);
		sequential_prefix_scan(d, states, popD, p3, ls, ls, popH->r[2].end, PUPA);
		sequential_prefix_scan(d, states, popD, p1, 0, popH->r[2].start, popH->r[3].end, IMMATURE);
		sequential_prefix_scan(d, states, popD, p2, 0, popH->r[3].start, popH->r[4].end, MATESEEKING);
		sequential_prefix_scan(d, states, popD, p3, ls, popH->r[4].start, popH->r[7].end, BMS);
		sequential_prefix_scan(d, states, popD, p1, 0, popH->r[5].start, popH->r[6].end, BMD);
		sequential_prefix_scan(d, states, popD, p2, 0, popH->r[6].start, popH->r[7].end, GRAVID);
	}
	global = ((popH->r[7].end + LOCAL_SIZE) / LOCAL_SIZE) * LOCAL_SIZE;
	icl_run_kernel(d->oldToNewAgents, 1, &global, &local, NULL, NULL, 11, (size_t)0, (void *)agents, (size_t)0,
	               (void *)ages, (size_t)0, (void *)states, (size_t)0, (void *)nagents, (size_t)0, (void *)nages,
	               (size_t)0, (void *)nstates, (size_t)0, (void *)p1, (size_t)0, (void *)p2, (size_t)0, (void *)p3,
	               (size_t)0, (void *)popD, (size_t)0, (void *)enbD);
}

static void swap(icl_buffer **a, icl_buffer **b)
{
	icl_buffer *t = *a;
	*a = *b;
	*b = t;
}

int sampo_abms_run(const sampo_abms_params *p, sampo_abms_result *res)
{
	if (!p || !res || !p->temp || !p->env || p->steps == 0 || p->max_agents < 2 * LOCAL_SIZE) return SAMPO_E_ARG;
	if (p->mode != SAMPO_MODE_REFERENCE && p->mode != SAMPO_MODE_CONFIG) return SAMPO_E_ARG;
	if (p->mode == SAMPO_MODE_CONFIG) {
		const sampo_layout *L = p->init_layout;
		if (!L || !p->init_agents || L->start[0] != 0) return SAMPO_E_ARG;
		for (int s = 0; s < 8; ++s)
			if (L->end[s] < L->start[s] || (s > 0 && L->start[s] <= L->end[s - 1])) return SAMPO_E_ARG;
		if ((uint64_t)L->end[7] + 8u * LOCAL_SIZE > p->max_agents) return SAMPO_E_ARG;
	}
	Driver D;
	memset(&D, 0, sizeof(D));
	D.p = p;
	{ /* readArguments, src/abms.c:557-583: the species header goes into every kernel's build options */
		const char *sp = p->species && *p->species ? p->species : "properties.h";
		size_t n = strlen(sp), k = strlen("properties.h");
		sampo_species check;
		if (strlen(sp) > 1024) return SAMPO_E_ARG;
		if ((n < k || strcmp(sp + n - k, "properties.h") != 0) && sampo_species_read(sp, &check)) return SAMPO_E_ARG;
		snprintf(D.opts, sizeof D.opts, "-Iinclude -DSPECIES=%s", sp);
	}
	D.maxAgents = p->max_agents;
	for (int c = 0; c < 4; ++c) D.hmr[c] = sm_hmr(p->dmr[c]);
	sc_glibc_seed(&D.rng, p->seed);
	res->status = 0;
	res->steps_done = 0;

	icl_init_devices(ICL_GPU);
	if (icl_get_num_devices() == 0) return SAMPO_E_HIP;
	D.dev = icl_get_device(0);
	icl_device *dev = D.dev;
	icl_timer *total = icl_init_timer(ICL_MILLI);
	icl_start_timer(total);

	/* main(), src/abms.c:606-672 */
	const cl_uint eggs0 = p->mode == SAMPO_MODE_CONFIG ? 0u : p->n_eggs;
	Enb enbH = {(cl_int)eggs0, eggs0};
	Bnc bncH = {0, 0, 0, 0};
	icl_buffer *seedsD = icl_create_buffer(dev, CL_MEM_READ_ONLY, NSEEDS * sizeof(Seeds));
	Seeds seedsH[NSEEDS];
	icl_buffer *enbD = icl_create_buffer(dev, CL_MEM_READ_WRITE, sizeof(Enb));
	icl_buffer *bnc = icl_create_buffer(dev, CL_MEM_READ_WRITE, sizeof(Bnc));
	size_t cap = D.maxAgents;
	icl_buffer *agents = icl_create_buffer(dev, CL_MEM_READ_WRITE, cap * 24);
	icl_buffer *ages = icl_create_buffer(dev, CL_MEM_READ_WRITE, cap * 16);
	icl_buffer *states = icl_create_buffer(dev, CL_MEM_READ_WRITE, cap * 8);
	icl_write_buffer(enbD, CL_TRUE, sizeof(Enb), &enbH, NULL, NULL);
	icl_write_buffer(bnc, CL_TRUE, sizeof(Bnc), &bncH, NULL, NULL);
	generate_seeds(&D, seedsD, seedsH);
	Population popH;
	memset(&popH, 0, sizeof(popH));
	create_initial_population(&D, agents, ages, states, &popH, enbD, seedsD);

	/* run(), src/abms.c:464-555 */
	icl_buffer *nagents = icl_create_buffer(dev, CL_MEM_READ_WRITE, cap * 24);
	icl_buffer *nages = icl_create_buffer(dev, CL_MEM_READ_WRITE, cap * 16);
	icl_buffer *nstates = icl_create_buffer(dev, CL_MEM_READ_WRITE, cap * 8);
	icl_buffer *popD = icl_create_buffer(dev, CL_MEM_READ_WRITE, 2 * sizeof(Population));
	size_t buff_size = cap > 128 ? cap : 128;
	icl_buffer *buff = icl_create_buffer(dev, CL_MEM_READ_WRITE, (buff_size + 2) * sizeof(cl_uint));
	icl_buffer *p1 = icl_create_buffer(dev, CL_MEM_READ_WRITE, (cap + 2) * sizeof(cl_int));
	icl_buffer *p2 = icl_create_buffer(dev, CL_MEM_READ_WRITE, (cap + 2) * sizeof(cl_int));
	icl_buffer *p3 = icl_create_buffer(dev, CL_MEM_READ_WRITE, (cap + 2) * sizeof(cl_int));
	icl_write_buffer_offset(popD, CL_TRUE, sizeof(Population), sizeof(Population), &popH, NULL, NULL);
	const char *o = D.opts;
	D.calcL1dePerGroup = icl_create_kernel(dev, "kernel/calcL1de.cl", "calcL1dePerGroup", o, ICL_SOURCE);
	D.calcL1deTotal = icl_create_kernel(dev, "kernel/calcL1de.cl", "calcL1deTotal", o, ICL_SOURCE);
	D.calcFemalesPerGroup = icl_create_kernel(dev, "kernel/calcGender.cl", "calcFemalesPerGroup", o, ICL_SOURCE);
	D.calcFemalesTotal = icl_create_kernel(dev, "kernel/calcGender.cl", "calcFemalesTotal", o, ICL_SOURCE);
	D.killAgents = icl_create_kernel(dev, "kernel/killAgents.cl",
	                                 p->mode == SAMPO_MODE_CONFIG ? "killAgentsStaged" : "killAgents", o, ICL_SOURCE);
	D.updateAgents = icl_create_kernel(dev, "kernel/updateAgents.cl", "updateAgents", o, ICL_SOURCE);
	D.oldToNewAgents = icl_create_kernel(dev, "kernel/oldToNewAgents.cl", "oldToNewAgents", o, ICL_SOURCE);
	D.serialScan = icl_create_kernel(dev, "kernel/serialScan.cl", "serialScan", o, ICL_SOURCE);
	D.preScan = icl_create_kernel(dev, "kernel/preScan.cl", "preScan", o, ICL_SOURCE);
	segmented_scan_init(2 * LOCAL_SIZE, D.maxAgents, dev, o, ICL_SOURCE);

	icl_timer *loop = icl_init_timer(ICL_MILLI);
	icl_start_timer(loop);
	int status = 0;
	for (cl_uint step = 0; step < p->steps; ++step) {
		sampo_row row;
		memset(&row, 0, sizeof(row));
		status = calc_stats(&D, ages, popD, buff, &popH, &row);
		if (status) break;
		Bnc b;
		icl_read_buffer(bnc, CL_TRUE, sizeof(Bnc), &b, NULL, NULL);
		row.numCycles = b.numCycles;
		row.sumCycles = b.sumCycles;
		row.bites = b.numBites;
		row.infectBites = b.numInfectBites;
		if (res->rows) res->rows[step] = row;
		res->agent_steps += (uint64_t)row.count[0] + row.count[1] + row.count[2] + row.count[3] + row.count[4] +
		                    row.count[5] + row.count[6] + row.count[7];
		generate_seeds(&D, seedsD, seedsH);
		update(&D, agents, ages, states, enbD, bnc, popD, seedsD, (cl_uint)((float)step * 1.0f) % 24u, &popH, step);
		/* new eggs are written from index 0 whatever their number: guard the buffers */
		Enb e;
		icl_read_buffer(enbD, CL_TRUE, sizeof(Enb), &e, NULL, NULL);
		if ((uint64_t)e.newEggs + popH.r[7].end + 8u * LOCAL_SIZE > D.maxAgents) {
			status = SAMPO_CAPACITY;
			break;
		}
		create_new_agents(&D, agents, ages, states, nagents, nages, nstates, enbD, popD, seedsD, p1, p2, p3, buff,
		                  &popH, step);
		swap(&agents, &nagents);
		swap(&ages, &nages);
		swap(&states, &nstates);
		res->steps_done = step + 1;
	}
	icl_finish(dev);
	res->loop_ms = icl_stop_timer(loop);
	res->status = status;
	Population fin;
	icl_read_buffer_offset(popD, CL_TRUE, sizeof(Population), sizeof(Population), &fin, NULL, NULL);
	for (int s = 0; s < 8; ++s) {
		res->layout.start[s] = fin.r[s].start;
		res->layout.end[s] = fin.r[s].end;
	}
	segmented_scan_release();

	/* storeHistogram, src/abms.c:160-183 */
	icl_kernel *hk = icl_create_kernel(dev, "kernel/ageHistogram.cl", "ageHistogram", D.opts, ICL_SOURCE);
	cl_uint zero[100];
	memset(zero, 0, sizeof(zero));
	icl_write_buffer(buff, CL_TRUE, 100 * sizeof(cl_uint), zero, NULL, NULL);
	size_t local = LOCAL_SIZE, global = ((D.maxAgents + LOCAL_SIZE) / LOCAL_SIZE) * LOCAL_SIZE;
	icl_run_kernel(hk, 1, &global, &local, NULL, NULL, 3, (size_t)0, (void *)ages, (size_t)0, (void *)popD,
	               (size_t)0, (void *)buff);
	icl_read_buffer(buff, CL_TRUE, 100 * sizeof(cl_uint), res->histogram, NULL, NULL);
	icl_release_kernel(hk);

	size_t span = fin.r[7].end;
	if (res->agents && span <= res->span_capacity) {
		icl_read_buffer(agents, CL_TRUE, span * 24, res->agents, NULL, NULL);
		icl_read_buffer(ages, CL_TRUE, span * 16, res->ages, NULL, NULL);
		icl_read_buffer(states, CL_TRUE, span * 8, res->states, NULL, NULL);
	}

	icl_release_kernel(D.calcL1dePerGroup);
	icl_release_kernel(D.calcL1deTotal);
	icl_release_kernel(D.calcFemalesPerGroup);
	icl_release_kernel(D.calcFemalesTotal);
	icl_release_kernel(D.killAgents);
	icl_release_kernel(D.updateAgents);
	icl_release_kernel(D.createEggs);
	icl_release_kernel(D.oldToNewAgents);
	icl_release_kernel(D.serialScan);
	icl_release_kernel(D.preScan);
	icl_release_buffers(6, nagents, nages, nstates, popD, buff, seedsD);
	icl_release_buffers(3, p1, p2, p3);
	icl_release_buffers(5, enbD, bnc, agents, ages, states);
	icl_release_timer(loop);
	res->total_ms = icl_stop_timer(total);
	icl_release_timer(total);
	icl_release_devices();
	return 0;
}
