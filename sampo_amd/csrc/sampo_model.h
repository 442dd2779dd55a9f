/*
 * sampo_model.h -- the agent model as the fused engine evaluates it, shared
 * by the device kernels and the host-side preparation of iteration 0.
 *
 * Storage is structure-of-arrays, 8 dwords per agent (32 bytes):
 *   AGE  ageInHours            (f32)     AUX   ovipositionAttemps | cumulativeLarvalDelay bits
 *   HIS  hoursInState          (f32)     AVAIL availableEggs      (u32)
 *   DLY  delay                 (f32)     CYC   cycleLength        (u32)
 *   CSD  cumulativeSporogonicDevelopment MISC  humanBloodmealCount (bits 0-15),
 *                                              numEggBatches (16-30), isFemale (31)
 * The agent's state is implied by the range it is stored in and `dead` lives
 * only in registers, so neither is stored.
 *
 * Index spaces of one instance at the start of an iteration:
 *   dense k      0..n-1, the ranges egg..gravid concatenated without gaps;
 *   virtual g    the reference's padded index (ranges start at multiples of
 *                64 after a gap of >= 1, kernel/oldToNewAgents.cl:131-160),
 *                which keys the random numbers: g = k + voff[s];
 *   physical p   slot in the SoA buffer: p = k + off_in[s].
 * The first `enew` eggs are the newborns of the previous iteration.  They are
 * never stored: every field of a newborn is a function of its index and the
 * seeds and temperature of the iteration that laid it (kernel/initAgents.cl:
 * 26-65), so the pass regenerates them instead of reading them.
 */
#ifndef SAMPO_MODEL_H
#define SAMPO_MODEL_H

#include "sampo_canon.h"
#include "../../include/sampo_engine.h"

enum { F_AGE = 0, F_HIS, F_DLY, F_AUX, F_CSD, F_AVAIL, F_CYC, F_MISC, NFIELD };
enum { S_EGG = 0, S_LARVA, S_PUPA, S_IMM, S_MS, S_BMS, S_BMD, S_GRAVID, S_DEAD };

#define SAMPO_PAD 64u  /* LOCAL_SIZE of the reference layout (include/agent.h:16) */

/* Parameters of one iteration of one instance. */
typedef struct {
	uint32_t step;
	uint32_t active;
	uint32_t cur;          /* buffer holding the input */
	uint32_t n;            /* agents alive at the start (dense count) */
	uint32_t kstart[9];
	uint32_t off_in[8];
	uint32_t voff[8];
	uint32_t pout[8];      /* output base of each new state in buffer cur^1 */
	uint32_t vstart[8];
	uint32_t cnt[8];
	uint32_t enew;         /* newborn eggs at the head of the egg range */
	uint32_t nb_set;       /* seed set those newborns were laid with */
	float nb_temp;         /* and the temperature */
	uint32_t tb;           /* totalBiomass (killAgents.cl:42) */
	uint32_t h;            /* worldTime */
	float temp;
	sampo_env env;
	uint32_t seed_set;     /* seed set of this iteration */
	float hmr_egg, hmr_pupa;
	float hmr_larva[100];
} sampo_iter;

typedef struct {
	uint32_t step, cur;
	uint32_t cnt[8], vstart[8], phys_in[8];
	uint32_t enew;
	uint32_t l1de, females, potinf;
	uint32_t ncyc, scyc, bites, ibites;
	float temp_t, temp_prev;
	sampo_env env_t;
} sampo_prep_in;

/* properties.h:29-31 and 39-60 */
SC_FN float sm_hatching(float p)
{
	float m, b;
	if (p < 0.5f) { m = 48.0f; b = 0.0f; }
	else if (p < 0.85f) { m = 68.5714f; b = -10.2857f; }
	else if (p < 0.90f) { m = 480.0f; b = -360.0f; }
	else if (p < 0.94f) { m = 600.00f; b = -468.0f; }
	else { m = 2400.0f; b = -2160.0f; }
	return (float)sc_f2u(m * p + b + 0.5f);
}

SC_FN float sm_egg_delay(float t, float p) { return (-0.923f * t + 60.923f) + sm_hatching(p); }

/* killAgents.cl:91-93: HMR = 1 - (1 - DMR)^(1/24) */
SC_FN float sm_hmr(float dmr) { return 1.0f - sc_powf(1.0f - dmr, 1.0f / 24.0f); }

/* killAgents.cl:85-101, adults by age in days */
SC_FN float sm_hmr_adult(uint32_t d)
{
	const float a = 0.1f, B = 25.0f, s = 0.1f;
	float e = sc_expf((float)d / B);
	return sm_hmr((a * e) / (1.0f + (a * B * s * (e - 1.0f))));
}

/* killAgents.cl:117-133, density-dependent larvae */
SC_FN float sm_hmr_larva(uint32_t d, uint32_t l1de, float cc)
{
	float dmr = 0.1f * sc_expf((float)l1de / ((float)d * cc * 1.0f));
	return sm_hmr(sc_minf(dmr, 0.8f));
}

SC_FN uint32_t sm_roundup(uint32_t v, uint32_t m) { return ((v + m - 1u) / m) * m; }

/* The calcStats/begin-of-iteration bookkeeping of src/abms.c:230-310 plus the
 * launch geometry of the fused pass.  Fills *o (all but hmr_larva), the
 * printStats row, and returns the instance status. */
SC_FN uint32_t sm_prep(uint32_t mode, uint32_t horizon, uint64_t max_agents, uint32_t phys,
                       const float hmr_stage[4], const sampo_prep_in *in, sampo_iter *o,
                       sampo_row *row)
{
	uint32_t n = 0;
	o->step = in->step;
	o->cur = in->cur;
	o->enew = in->enew;
	for (int s = 0; s < 8; ++s) {
		o->kstart[s] = n;
		n += in->cnt[s];
		o->cnt[s] = in->cnt[s];
		o->vstart[s] = in->vstart[s];
	}
	o->kstart[8] = n;
	o->n = n;
	for (int s = 0; s < 8; ++s) {
		o->voff[s] = in->vstart[s] - o->kstart[s];
		o->off_in[s] = in->phys_in[s] - o->kstart[s] - (s == 0 ? in->enew : 0u);
	}
	/* an agent moves at most one state per iteration (plus gravid -> BMS), so
	 * these bound the output size of each state */
	const uint32_t *c = in->cnt;
	uint64_t bound[8] = {c[0], (uint64_t)c[1] + c[0], (uint64_t)c[2] + c[1], (uint64_t)c[3] + c[2],
	                     (uint64_t)c[4] + c[3], (uint64_t)c[5] + c[4] + c[7], (uint64_t)c[6] + c[5],
	                     (uint64_t)c[7] + c[6]};
	uint64_t pos = 0;
	for (int s = 0; s < 8; ++s) {
		o->pout[s] = (uint32_t)pos;
		pos += ((bound[s] + SAMPO_PAD - 1u) / SAMPO_PAD) * SAMPO_PAD;
	}
	for (int s = 0; s < 8; ++s) row->count[s] = c[s];
	uint32_t adults_range = (in->vstart[7] + c[7]) - in->vstart[3];
	row->l1de = c[1] > 0 ? in->l1de : 0u;
	row->females = adults_range > 0 ? in->females : 0u;
	row->potInfective = adults_range > 0 ? in->potinf : 0u;
	row->numCycles = in->ncyc;
	row->sumCycles = in->scyc;
	row->bites = in->bites;
	row->infectBites = in->ibites;
	uint32_t status;
	row->valid = 1u;
	if (in->step >= horizon) {
		status = SAMPO_DONE;
	} else if (!((double)n <= (double)max_agents * 0.95)) {
		status = SAMPO_CAPACITY;
		row->valid = 0u;
	} else if (n == 0u) {
		status = SAMPO_EXTINCT;
		row->valid = 0u;
	} else if (pos > (uint64_t)phys) {
		status = SAMPO_PHYS_FULL;
	} else {
		status = SAMPO_RUNNING;
	}
	o->active = status == SAMPO_RUNNING;
	o->tb = in->l1de + (in->vstart[0] + c[0]) + c[2];
	o->h = sc_f2u((float)in->step * 1.0f) % 24u;
	o->temp = in->temp_t;
	o->env = in->env_t;
	o->seed_set = in->step + 1u;
	o->nb_set = in->step;
	o->nb_temp = in->temp_prev;
	if (mode == SAMPO_MODE_CONFIG) {
		o->hmr_egg = hmr_stage[0];
		o->hmr_pupa = hmr_stage[2];
	} else {
		o->hmr_egg = sm_hmr(0.1f);
		o->hmr_pupa = sm_hmr(0.1f);
	}
	return status;
}

#endif
