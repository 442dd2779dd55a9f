/*
 * sampo_model.h -- the agent model as the fused engine evaluates it, shared
 * by the device kernels and the host-side preparation of iteration 0.
 *
 * Storage is structure-of-arrays, 7 dwords per agent (28 bytes):
 *   AH   ageInHours (bits 0-15) and hoursInState (16-31) as integers
 *   DLY  delay                             (f32)
 *   AUX  ovipositionAttemps | cumulativeLarvalDelay bits
 *   CSD  cumulativeSporogonicDevelopment    (f32)
 *   AVAIL availableEggs                    (u32)
 *   CH   cycleLength (bits 0-15), humanBloodmealCount (16-31)
 *   NF   numEggBatches (bits 0-14), isFemale (15)
 * The two hour counts are whole numbers of hours: they start at 0 or 24
 * and grow by the step of one hour (src/abms.c:15), so each is held exactly
 * in 16 bits while it stays below 2^16; cycleLength grows by at most one an
 * iteration.  Admission refuses initial populations and horizons that could
 * take any of them past their field (sampo_engine_admit).  The agent's state
 * is implied by the range it is stored in and `dead` lives only in
 * registers, so neither is stored.
 *
 * Words that are zero for every agent of a state are not stored for it
 * (sm_fields): an egg keeps AH, DLY and NF, a larva or pupa AH, DLY, AUX and
 * NF, an adult all seven.  This holds because the update rules
 * (kernel/updateAgents.cl:39-270) only set cumulativeSporogonicDevelopment,
 * availableEggs, cycleLength, humanBloodmealCount and numEggBatches in adult
 * states (the development sum grows only once humanBloodmealCount > 0, and
 * only by a finite amount at a finite temperature), and an egg's
 * cumulativeLarvalDelay is the zero initAgents.cl:53 gives it.  An instance
 * whose initial population or temperature series does not meet these
 * conditions stores all seven words in every state (Inst.compact = 0).
 *
 * Index spaces of one instance at the start of an iteration:
 *   dense k      0..n-1, the states egg..gravid concatenated without gaps;
 *   virtual g    the reference's padded index (ranges start at multiples of
 *                64 after a gap of >= 1, kernel/oldToNewAgents.cl:131-160),
 *                which keys the random numbers;
 *   physical p   slot in the SoA buffer.
 * The agents of one iteration are a list of segments in dense order: runs
 * of one state stored contiguously, each with its dense start, physical base
 * and virtual base.  The first segment may hold the `enew` newborn eggs of
 * the previous iteration.  They are never stored: every field of a newborn
 * is a function of its index and the seeds and temperature of the iteration
 * that laid it (kernel/initAgents.cl:26-65), so the pass regenerates them.
 */
#ifndef SAMPO_MODEL_H
#define SAMPO_MODEL_H

#include "sampo_canon.h"
#include "../../include/sampo_engine.h"

enum { F_AH = 0, F_DLY, F_AUX, F_CSD, F_AVAIL, F_CH, F_NF, NFIELD };
enum { S_EGG = 0, S_LARVA, S_PUPA, S_IMM, S_MS, S_BMS, S_BMD, S_GRAVID, S_DEAD };

#define SM_ALL_FIELDS ((1u << NFIELD) - 1u)

/* the stored words of state s: a bit per word (the others are zero) */
SC_FN uint32_t sm_fields(uint32_t s, uint32_t compact)
{
	if (!compact || s > S_PUPA) return SM_ALL_FIELDS;
	const uint32_t imm = (1u << F_AH) | (1u << F_DLY) | (1u << F_NF);
	return s == S_EGG ? imm : imm | (1u << F_AUX);
}

#define SAMPO_PAD 64u  /* LOCAL_SIZE of the reference layout (include/agent.h:16) */

/* The species functions of include/properties.h, evaluated from a species
 * table (sampo_species) in the arithmetic the header uses. */

/* larvaDevelopment (properties.h:105-109): T clamped to the table's range,
 * truncated, minus the range start; the index is clamped to the 32 entries
 * (the reference reads past the table for T >= 44 degC) */
SC_FN int sm_dev_index(const sampo_species *sp, float t)
{
	float tc = sc_minf(sc_maxf(t, sp->dev_temp_range[0]), sp->dev_temp_range[1]);
	int ix = (int)tc - (int)sp->dev_temp_range[0];
	return ix > 31 ? 31 : ix;
}

SC_FN float sm_dev_rate(const sampo_species *sp, float t) { return sp->dev_rates[sm_dev_index(sp, t)]; }

/* eggsHatching (properties.h:39-60): the first band whose threshold exceeds
 * p, rounded to an integer number of hours.  The thresholds ascend
 * (sampo_species_check), so that band's index is the number of thresholds p
 * is not below: counted without branches, it costs one round of loads. */
SC_FN float sm_hatching(const float *thr, const float *m, const float *b, float p)
{
	const int i = (int)!(p < thr[0]) + (int)!(p < thr[1]) + (int)!(p < thr[2]) + (int)!(p < thr[3]);
	return (float)sc_f2u(m[i] * p + b[i] + 0.5f);
}

SC_FN float sm_linear(const float *ab, float t) { return ab[0] * t + ab[1]; }

/* initAgents.cl:57: eggsIncubation(T) + eggsHatching(p) */
SC_FN float sm_egg_delay(const sampo_species *sp, float t, float p)
{
	return sm_linear(sp->egg_incubation, t) +
	       sm_hatching(sp->hatch_threshold, sp->hatch_slope, sp->hatch_offset, p);
}

/* pupaDelay (properties.h:133-135): binary64, rounded once */
SC_FN float sm_pupa_delay(const sampo_species *sp, float t)
{
	return (float)(sp->pupa_delay[0] * (double)t + sp->pupa_delay[1]);
}

/* the transition-time predicate of state s (7: egg laying) at hour h */
SC_FN int sm_open(const sampo_species *sp, int s, uint32_t h) { return (int)((sp->window[s] >> h) & 1u); }

/* generateEggs (properties.h:257-266) for batch nb from the pair (u0, u1) */
SC_FN uint32_t sm_egg_batch(float u0, float u1, float mean, float sd, float decay, uint32_t nb)
{
	uint32_t eggs = sc_f2u(sc_box_muller(u0, u1, mean, sd) + 0.5f);
	return sc_f2u((float)eggs * sc_powf(decay, (float)nb) + 0.5f);
}

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

/* HybridTaus (sc_hybrid_taus_bits, include/gpuRand.h:13-50) with the work
 * that depends on the seeds alone done once per set.  Of a set (z1, z2, z3,
 * z4), the first and third Tausworthe components do not involve the
 * work-item index gid, the second steps z2 + 128 + gid, and the LCG term is
 * 1664525 * (z4 * gid) + 1013904223 = (1664525 * z4) * gid + 1013904223
 * modulo 2^32.  A prepared set holds {the first and third components xored,
 * z2 + 128, 1664525 * z4, 0}; a draw then costs one Tausworthe step, one
 * multiply-add and two xors instead of three steps and two multiplies. */
SC_FN void sm_taus_prepare(const uint32_t z[4], uint32_t out[4])
{
	const uint32_t k = sc_taus_step(z[0] + 128u, 13, 19, 12, 0xfffffffeu) ^
	                   sc_taus_step(z[2] + 128u, 3, 11, 17, 0xfffffff0u);
	const uint32_t b = z[1] + 128u, m = 1664525u * z[3];
	out[0] = k;
	out[1] = b;
	out[2] = m;
	out[3] = 0u;
}

/* sc_hybrid_taus(z1, z2, z3, z4, gid) of the set p prepared by sm_taus_prepare */
SC_FN float sm_taus_draw(const uint32_t *p, uint32_t gid)
{
	const uint32_t bits = p[0] ^ sc_taus_step(p[1] + gid, 2, 25, 4, 0xfffffff8u) ^ (p[2] * gid + 1013904223u);
	return sc_float(0x2f800000u) * (float)bits;
}

#endif
