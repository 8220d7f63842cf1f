/*
 * cli_util.c -- helpers shared by the command-line tools: option value
 * parsing, debug output (misc/debug.c levels: 0 error .. 4 debug), image
 * size estimation, pattern estimation and data scaling as BART computes
 * them.
 */
#include "tools.h"
#include "../../include/bart_amd.h"
#include "cfl.h"

#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int dlevel = 2;  /* DP_INFO */

int debug_level(void)
{
	return dlevel;
}

void set_debug_level(int l)
{
	dlevel = l;
}

void debug_printf(int level, const char* fmt, ...)
{
	if (level > dlevel)
		return;

	va_list ap;
	va_start(ap, fmt);
	vfprintf(stderr, fmt, ap);
	va_end(ap);
}

double timestamp(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void parse_vec3(const char* s, long v[3], const char* opt)
{
	char* end = NULL;
	const char* p = s;

	for (int i = 0; i < 3; i++) {

		v[i] = strtol(p, &end, 10);

		if (end == p)
			cli_error("option %s: expected x:y:z, got '%s'", opt, s);

		if (2 == i)
			break;

		if (':' != *end)
			cli_error("option %s: expected x:y:z, got '%s'", opt, s);

		p = end + 1;
	}

	if ('\0' != *end)
		cli_error("option %s: expected x:y:z, got '%s'", opt, s);
}

float parse_float(const char* s, const char* opt)
{
	char* end = NULL;
	float f = strtof(s, &end);

	if (end == s || '\0' != *end)
		cli_error("option %s: expected a number, got '%s'", opt, s);

	return f;
}

int parse_int(const char* s, const char* opt)
{
	char* end = NULL;
	long l = strtol(s, &end, 10);

	if (end == s || '\0' != *end)
		cli_error("option %s: expected an integer, got '%s'", opt, s);

	return (int)l;
}

void unsupported_option(const char* tool, const char* opt)
{
	cli_error("%s: option %s is not supported (GPU NUFFT / CG-SENSE path only)", tool, opt);
}

void parse_nufft_conf(const char* arg, struct nufft_conf_s* conf)
{
	char buf[512];
	snprintf(buf, sizeof(buf), "%s", arg);

	for (char* tok = strtok(buf, ","); NULL != tok; tok = strtok(NULL, ",")) {

		char* val = strchr(tok, '=');

		if (NULL != val)
			*val++ = '\0';

		if (!strcmp(tok, "r") || !strcmp(tok, "no-toeplitz"))
			conf->toeplitz = false;
		else if (!strcmp(tok, "P") || !strcmp(tok, "periodic"))
			conf->periodic = true;
		else if (!strcmp(tok, "1"))
			conf->decomp = false;
		else if (!strcmp(tok, "lowmem"))
			conf->lowmem = true;
		else if (!strcmp(tok, "zero-mem"))
			conf->zero_overhead = true;
		else if (!strcmp(tok, "no-precomp"))
			conf->precomp = false;
		else if ((!strcmp(tok, "o") || !strcmp(tok, "oversampling")) && NULL != val)
			conf->os = parse_float(val, "--nufft-conf oversampling");
		else if ((!strcmp(tok, "w") || !strcmp(tok, "width")) && NULL != val)
			conf->width = parse_float(val, "--nufft-conf width");
		else if (!strcmp(tok, "real-psf"))
			conf->real = true;
		else if (!strcmp(tok, "compress-psf"))
			conf->compress_psf = true;
		else if (!strcmp(tok, "decomposed-psf"))
			conf->decomposed_psf = true;
		else if (!strcmp(tok, "upper-triag-psf"))
			conf->upper_triag = true;
		else
			cli_error("Sub-option could not be parsed: %s", tok);
	}
}

void estimate_im_dims(const long traj_dims[DIMS], const complex float* traj, long dims[3])
{
	int T = (int)traj_dims[0];
	float mx[3] = { 0.f, 0.f, 0.f };
	long n = cfl_size(DIMS - 1, traj_dims + 1);

	if (T > 3)
		cli_error("trajectory with %d coordinates", T);

	for (long i = 0; i < n; i++)
		for (int j = 0; j < T; j++)
			mx[j] = fmaxf(cabsf(traj[j + T * i]), mx[j]);

	for (int j = 0; j < 3; j++)
		dims[j] = (j >= T || 0.f == mx[j]) ? 1 : (long)(2.f * ceilf(mx[j]));
}

void estimate_pattern(const long ksp_dims[DIMS], complex float* pattern, const complex float* ksp)
{
	long inner = cfl_size(COIL_DIM, ksp_dims);
	long C = ksp_dims[COIL_DIM];
	long outer = cfl_size(DIMS - COIL_DIM - 1, ksp_dims + COIL_DIM + 1);

	for (long o = 0; o < outer; o++) {
		for (long i = 0; i < inner; i++) {

			float rss = 0.f;

			for (long c = 0; c < C; c++) {
				complex float v = ksp[i + inner * (c + C * o)];
				rss += crealf(v * conjf(v));
			}

			pattern[i + inner * o] = (0.f == sqrtf(rss)) ? 0.f : 1.f;
		}
	}
}

static int cmp_abs(const void* _a, const void* _b)
{
	const complex float* a = _a;
	const complex float* b = _b;
	return (int)copysignf(1.f, cabsf(*a) - cabsf(*b));
}

float estimate_scaling_norm(long n, complex float* x)
{
	qsort(x, (size_t)n, sizeof(complex float), cmp_abs);

	float median = cabsf(x[n / 2]);
	float p90 = cabsf(x[(long)trunc((double)n * 0.9)]);
	float max = cabsf(x[n - 1]);

	float scale = ((max - p90) < 2 * (p90 - median)) ? p90 : max;

	debug_printf(3, "Scaling: %f%c (max = %f/p90 = %f/median = %f)\n", scale, (scale == max) ? '!' : ' ', max, p90,
		     median);

	return scale;
}
