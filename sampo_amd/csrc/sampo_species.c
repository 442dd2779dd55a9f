/*
 * sampo_species.c -- species tables: the include/properties.h defaults, their
 * check, and the species file reader.  Host C, linked into both libraries.
 */
#include <ctype.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sampo_engine.h"

#define ALL_HOURS 0xffffffu
#define NIGHT 0xfc007fu /* worldTime <= 6u || worldTime >= 18u */

void sampo_species_default(sampo_species *o)
{
	static const sampo_species d = {
		.egg_incubation = {-0.923f, 60.923f},                                /* properties.h:29-31 */
		.hatch_threshold = {0.5f, 0.85f, 0.90f, 0.94f},                      /* properties.h:39-60 */
		.hatch_slope = {48.0f, 68.5714f, 480.0f, 600.00f, 2400.0f},
		.hatch_offset = {0.0f, -10.2857f, -360.0f, -468.0f, -2160.0f},
		.larva_delay = {1.0f, 0.1f},                                         /* properties.h:87-91 */
		.dev_temp_range = {12.0f, 48.0f},                                    /* properties.h:105-109 */
		.dev_rates = {0.000175044f, 0.000549094f, 0.00109031f, 0.001494158f, 0.001752487f, 0.001962866f,
		              0.002171238f, 0.002392989f, 0.002633689f, 0.00289619f, 0.003182685f, 0.003495287f,
		              0.003836204f, 0.004207789f, 0.004612566f, 0.005053243f, 0.005532729f, 0.006054144f,
		              0.006620836f, 0.007236393f, 0.007904657f, 0.008629745f, 0.009416051f, 0.010267388f,
		              0.011113704f, 0.011960019f, 0.012806335f, 0.010401163f, 0.007995991f, 0.00033845f,
		              0.00000533086f, 0.0000000840986f},                      /* properties.h:94-98 */
		.pupa_delay = {-0.923, 60.923},                                      /* properties.h:133-135 */
		.immature_delay = {-2.667f, 120.0f},                                 /* properties.h:159-161 */
		.mate_seeking_delay = 1.0f,                                          /* properties.h:185-187 */
		.bms_delay = 0.0f,                                                   /* properties.h:212-214 */
		.bmd_delay = {-1.231f, 77.231f},                                     /* properties.h:239-241 */
		.egg_batch = {170.0f, 30.0f, 0.8f},                                  /* properties.h:257-266 */
		.lay_probability = 0.25f,                                            /* updateAgents.cl:233 */
		.bite_human = 1.0f, .bite_indoor = 1.0f, .rest_indoor = 1.0f,        /* properties.h:12-14 */
		/* properties.h:68-70, 117-119, 143-145, 169-171, 195-197, 222-224, 249-251, 280-282 */
		.window = {ALL_HOURS, NIGHT, NIGHT, ALL_HOURS, 1u << 18, NIGHT, ALL_HOURS, NIGHT},
		.reserved = 0u,
	};
	*o = d;
}

static int prob_ok(float p) { return p >= 0.0f && p <= 1.0f; }

int sampo_species_check(const sampo_species *sp)
{
	if (!sp) return SAMPO_E_ARG;
	for (int i = 1; i < 4; ++i)
		if (!(sp->hatch_threshold[i - 1] <= sp->hatch_threshold[i])) return SAMPO_E_ARG;
	const float lo = sp->dev_temp_range[0], hi = sp->dev_temp_range[1];
	if (!(lo <= hi) || lo != (float)(int)lo || lo < -1000.0f || hi > 1000.0f) return SAMPO_E_ARG;
	if (!(sp->larva_delay[1] >= 0.0f) || !(sp->egg_batch[1] >= 0.0f) || !(sp->egg_batch[2] >= 0.0f))
		return SAMPO_E_ARG;
	if (!prob_ok(sp->lay_probability) || !prob_ok(sp->bite_human) || !prob_ok(sp->bite_indoor) ||
	    !prob_ok(sp->rest_indoor))
		return SAMPO_E_ARG;
	for (int s = 0; s < 8; ++s)
		if (sp->window[s] & ~ALL_HOURS) return SAMPO_E_ARG;
	if (sp->reserved) return SAMPO_E_ARG;
	return 0;
}

/* member name -> offset, count, kind (0 float, 1 double) */
typedef struct {
	const char *name;
	size_t off;
	int n, kind;
} Field;

#define F(m, n) {#m, offsetof(sampo_species, m), n, 0}
static const Field fields[] = {
	F(egg_incubation, 2), F(hatch_threshold, 4), F(hatch_slope, 5), F(hatch_offset, 5), F(larva_delay, 2),
	F(dev_temp_range, 2), F(dev_rates, 32), {"pupa_delay", offsetof(sampo_species, pupa_delay), 2, 1},
	F(immature_delay, 2), F(mate_seeking_delay, 1), F(bms_delay, 1), F(bmd_delay, 2), F(egg_batch, 3),
	F(lay_probability, 1), F(bite_human, 1), F(bite_indoor, 1), F(rest_indoor, 1),
};
#undef F
static const char *const windows[8] = {"window_egg", "window_larva", "window_pupa", "window_immature",
                                       "window_mate_seeking", "window_bms", "window_bmd", "window_lay"};

/* hours as "all", "none", "h" or "a-b" tokens */
static int parse_hours(char *rest, uint32_t *mask)
{
	uint32_t m = 0;
	for (char *tok = strtok(rest, " \t,\r\n"); tok; tok = strtok(NULL, " \t,\r\n")) {
		if (!strcmp(tok, "all")) { m |= ALL_HOURS; continue; }
		if (!strcmp(tok, "none")) continue;
		char *e;
		long a = strtol(tok, &e, 10), b = a;
		if (e == tok) return -1;
		if (*e == '-') {
			char *e2;
			b = strtol(e + 1, &e2, 10);
			if (e2 == e + 1) return -1;
			e = e2;
		}
		if (*e || a < 0 || b > 23 || a > b) return -1;
		for (long h = a; h <= b; ++h) m |= 1u << h;
	}
	*mask = m;
	return 0;
}

int sampo_species_read(const char *path, sampo_species *out)
{
	if (!path || !out) return SAMPO_E_ARG;
	FILE *f = fopen(path, "r");
	if (!f) return SAMPO_E_ARG;
	sampo_species sp;
	sampo_species_default(&sp);
	char line[4096];
	int rc = 0;
	while (!rc && fgets(line, sizeof line, f)) {
		char *hash = strchr(line, '#');
		if (hash) *hash = '\0';
		char *p = line;
		while (isspace((unsigned char)*p)) ++p;
		if (!*p) continue;
		char *name = p;
		while (*p && !isspace((unsigned char)*p)) ++p;
		if (*p) *p++ = '\0';
		int w = -1;
		for (int s = 0; s < 8; ++s)
			if (!strcmp(name, windows[s])) w = s;
		if (w >= 0) {
			if (parse_hours(p, &sp.window[w])) rc = SAMPO_E_ARG;
			continue;
		}
		const Field *fd = NULL;
		for (size_t i = 0; i < sizeof fields / sizeof fields[0]; ++i)
			if (!strcmp(name, fields[i].name)) fd = &fields[i];
		if (!fd) { rc = SAMPO_E_ARG; break; }
		char *base = (char *)&sp + fd->off;
		for (int i = 0; i < fd->n; ++i) {
			while (isspace((unsigned char)*p) || *p == ',') ++p;
			char *e;
			if (fd->kind) ((double *)base)[i] = strtod(p, &e);
			else ((float *)base)[i] = strtof(p, &e);
			if (e == p) { rc = SAMPO_E_ARG; break; }
			p = e;
			if (*p == 'f') ++p; /* C float suffix, as in the species header */
		}
		while (isspace((unsigned char)*p) || *p == ',') ++p;
		if (!rc && *p) rc = SAMPO_E_ARG; /* more values than the member holds */
	}
	fclose(f);
	if (!rc) rc = sampo_species_check(&sp);
	if (!rc) *out = sp;
	return rc;
}
