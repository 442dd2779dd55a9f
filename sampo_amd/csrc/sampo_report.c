/*
 * sampo_report.c -- text input and output of the reference driver around
 * the agent step (src/abms.c:62-137, 160-183), for the rows, ranges and
 * histogram the engine and the driver loop return as data.
 */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "../../include/sampo_report.h"

/* append printf output at *len; fails when buf would overflow */
static int put(char *buf, size_t cap, size_t *len, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	int n = vsnprintf(buf + *len, *len < cap ? cap - *len : 0, fmt, ap);
	va_end(ap);
	if (n < 0 || *len + (size_t)n >= cap) return SAMPO_E_ARG;
	*len += (size_t)n;
	return 0;
}

int sampo_format_stats(char *buf, size_t cap, uint32_t step, float hours_per_step, float temp,
                       const sampo_row *row)
{
	if (!buf || !row) return SAMPO_E_ARG;
	const uint32_t *c = row->count;
	/* Stats of calcStats (src/abms.c:287-302) from the row */
	const uint32_t mature = c[3] + c[4] + c[5] + c[6] + c[7];
	const uint32_t l1de = c[1] > 0u ? row->l1de : 0u;
	const uint32_t biomass = l1de + c[0] + c[2];
	const uint32_t males = mature - row->females;
	const uint32_t cycle = row->numCycles > 0u ? row->sumCycles / row->numCycles : 0u;
	size_t len = 0;
	int rc = put(buf, cap, &len, "%0.1f, %0.1f", (double)(hours_per_step * (float)step), (double)temp);
	if (!rc)
		rc = put(buf, cap, &len, ", %d, %d, %d, %d, %d, %d, %d, %d, %d", (int)mature, (int)c[3], (int)c[4], (int)c[5],
		         (int)c[6], (int)c[7], (int)row->females, (int)row->potInfective, (int)males);
	if (!rc) rc = put(buf, cap, &len, ", %d, %d, %d, %d, %d, ", (int)c[0], (int)c[1], (int)c[2], (int)l1de, (int)biomass);
	if (!rc) rc = put(buf, cap, &len, "%d, %d, %d\r\n", (int)cycle, (int)row->bites, (int)row->infectBites);
	return rc ? rc : (int)len;
}

int sampo_format_population(char *buf, size_t cap, const sampo_layout *layout)
{
	static const char *const names[8] = {"Eggs \t\t",       "Larvae \t\t", "Pupae \t\t", "Immatures \t",
	                                     "MateSeeking \t", "BMS \t\t",    "BMD \t\t",   "Gravids \t"};
	if (!buf || !layout) return SAMPO_E_ARG;
	size_t len = 0;
	for (int s = 0; s < 8; ++s) {
		const uint32_t a = layout->start[s], b = layout->end[s];
		int rc = put(buf, cap, &len, "%s[%d\t%d]\t%d\r\n", names[s], (int)a, (int)b, (int)(b - a));
		if (rc) return rc;
	}
	return (int)len;
}

int sampo_format_histogram(char *buf, size_t cap, const uint32_t hist[100])
{
	if (!buf || !hist) return SAMPO_E_ARG;
	size_t len = 0;
	for (int i = 0; i < 100; ++i) {
		int rc = put(buf, cap, &len, "%d, ", (int)hist[i]);
		if (rc) return rc;
	}
	return (int)len;
}

int sampo_read_single_val_file(const char *path, uint32_t n, float *out)
{
	if (!path || (!out && n)) return SAMPO_E_ARG;
	FILE *f = fopen(path, "r");
	if (!f) return SAMPO_E_ARG;
	uint32_t i = 0;
	while (i < n && fscanf(f, "%f", &out[i]) == 1) ++i;
	fclose(f);
	return (int)i;
}

int sampo_read_environment(const char *path, uint32_t n, float carrying_capacity, sampo_env *out)
{
	if (!path || (!out && n)) return SAMPO_E_ARG;
	FILE *f = fopen(path, "r");
	if (!f) return SAMPO_E_ARG;
	uint32_t i = 0;
	for (; i < n; ++i) {
		float cov[4], eff[4], success;
		int ok = 1;
		for (int k = 0; k < 4 && ok; ++k) ok = fscanf(f, "%f, %f, ", &cov[k], &eff[k]) == 2;
		if (!ok || fscanf(f, "%f", &success) != 1) break;
		out[i].IRSValue = cov[0] * eff[0];
		out[i].ITNValue = cov[1] * eff[1];
		out[i].larvacideValue = cov[2] * eff[2];
		out[i].oviTrapValue = cov[3] * eff[3];
		out[i].bloodmealSuccess = success;
		out[i].carryingCapacity = carrying_capacity;
	}
	fclose(f);
	return (int)i;
}
