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

enum { F_AGE = 0, F_HIS, F_DLY, F_AUX, F_CSD, F_AVAIL, F_CYC, F_MISC, NFIELD };
enum { S_EGG = 0, S_LARVA, S_PUPA, S_IMM, S_MS, S_BMS, S_BMD, S_GRAVID, S_DEAD };

#define SAMPO_PAD 64u  /* LOCAL_SIZE of the reference layout (include/agent.h:16) */

/* devRates, include/properties.h:94-98: hourly larval development at
 * 12, 13, ..., 43 degC; the one initializer every table in the library uses */
#define SM_DEV_RATES_INIT { \
	0.000175044f, 0.000549094f, 0.00109031f, 0.001494158f, 0.001752487f, 0.001962866f, \
	0.002171238f, 0.002392989f, 0.002633689f, 0.00289619f, 0.003182685f, 0.003495287f, \
	0.003836204f, 0.004207789f, 0.004612566f, 0.005053243f, 0.005532729f, 0.006054144f, \
	0.006620836f, 0.007236393f, 0.007904657f, 0.008629745f, 0.009416051f, 0.010267388f, \
	0.011113704f, 0.011960019f, 0.012806335f, 0.010401163f, 0.007995991f, 0.00033845f, \
	0.00000533086f, 0.0000000840986f}

/* devRates index of larvaDevelopment (properties.h:105-109): T clamped to
 * [12, 48], truncated, minus 12, and clamped to the 32-entry table */
SC_FN int sm_dev_index(float t)
{
	float tc = sc_minf(sc_maxf(t, 12.0f), 48.0f);
	int ix = (int)tc - 12;
	return ix > 31 ? 31 : ix;
}

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

#endif
