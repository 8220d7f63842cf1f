/*
 * cfl.c -- reading and writing BART's .hdr/.cfl and .ra array files.
 *
 * Layouts follow misc/io.c:380-468 (write_cfl_header), misc/io.c:560-650
 * (parse_cfl_header: "# Keyword" sections, dims until the end of the line,
 * dims past D must be 1) and misc/io.c:960-1040 (read_ra/write_ra).  Data
 * is read into memory instead of being mapped (misc/mmio.c maps it); the
 * bytes on disk are the same.
 */
#include "cfl.h"

#include <stdarg.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define RA_MAGIC 0x7961727261776172ULL
#define RA_TYPE_COMPLEX 4ULL

static char* command_line = NULL;

void cli_error(const char* fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	fprintf(stderr, "ERROR: ");
	vfprintf(stderr, fmt, ap);
	va_end(ap);
	size_t n = strlen(fmt);
	if (0 == n || '\n' != fmt[n - 1])
		fputc('\n', stderr);
	exit(1);
}

void cfl_set_command(int argc, char* argv[])
{
	size_t len = 1;

	for (int i = 0; i < argc; i++)
		len += strlen(argv[i]) + 1;

	free(command_line);
	command_line = calloc(len, 1);

	for (int i = 0; i < argc; i++) {
		if (i > 0)
			strcat(command_line, " ");
		strcat(command_line, argv[i]);
	}
}

long cfl_size(int D, const long dims[])
{
	long s = 1;

	for (int i = 0; i < D; i++)
		s *= dims[i];

	return s;
}

static bool has_suffix(const char* name, const char* suf)
{
	size_t n = strlen(name), m = strlen(suf);
	return (n > m) && (0 == strcmp(name + n - m, suf));
}

static void read_exact(FILE* f, void* buf, size_t n, const char* name)
{
	if (n != fread(buf, 1, n, f))
		cli_error("Loading file %s: short read", name);
}

static void write_exact(FILE* f, const void* buf, size_t n, const char* name)
{
	if (n != fwrite(buf, 1, n, f))
		cli_error("Creating file %s: short write", name);
}

// dims of a .hdr file (parse_cfl_header, misc/io.c:560-650)
static void parse_header(const char* text, const char* name, int D, long dims[])
{
	const char* p = text;
	bool ok = false;

	for (int i = 0; i < D; i++)
		dims[i] = 1;

	while (*p) {

		char keyword[32];
		int delta = 0;

		if (1 != sscanf(p, "# %31s\n%n", keyword, &delta) || 0 == delta)
			break;

		p += delta;

		if (0 == strcmp(keyword, "Dimensions")) {

			int i = 0;
			long val;

			while (1 == sscanf(p, "%ld%n", &val, &delta)) {

				// the dims end at the end of the line
				const char* q = p;
				bool eol = false;
				while (q < p + delta && (' ' == *q || '\t' == *q || '\n' == *q)) {
					if ('\n' == *q)
						eol = true;
					q++;
				}
				if (eol)
					break;

				p += delta;

				if (val < 0)
					cli_error("Loading cfl file %s: negative dimension", name);

				if (i < D)
					dims[i] = val;
				else if (1 != val)
					cli_error("Loading cfl file %s: more than %d dimensions", name, D);

				i++;
			}

			ok = true;
		}

		// skip to the next line starting with '#'
		const char* nl = strchr(p, '\n');

		while (NULL != nl && '#' != nl[1])
			nl = strchr(nl + 1, '\n');

		if (NULL == nl)
			break;

		p = nl + 1;
	}

	if (!ok)
		cli_error("Loading cfl file %s: no dimensions in header", name);
}

static complex float* load_ra(const char* name, int D, long dims[])
{
	FILE* f = fopen(name, "rb");

	if (NULL == f)
		cli_error("Loading ra file %s", name);

	uint64_t h[6];
	read_exact(f, h, sizeof(h), name);

	if (RA_MAGIC != h[0] || 0 != (h[1] & 1) || RA_TYPE_COMPLEX != h[2] || 8 != h[3] || h[5] > 100)
		cli_error("Loading ra file %s: not a little-endian complex float array", name);

	uint64_t nd = h[5];
	uint64_t rd[100];
	read_exact(f, rd, sizeof(uint64_t) * nd, name);

	for (int i = 0; i < D; i++)
		dims[i] = 1;

	for (uint64_t i = 0; i < nd; i++) {
		if ((int)i < D)
			dims[i] = (long)rd[i];
		else if (1 != rd[i])
			cli_error("Loading ra file %s: more than %d dimensions", name, D);
	}

	long n = cfl_size(D, dims);

	if ((uint64_t)n * 8 != h[4])
		cli_error("Loading ra file %s: size mismatch", name);

	complex float* data = malloc(sizeof(complex float) * (n > 0 ? n : 1));
	read_exact(f, data, sizeof(complex float) * n, name);
	fclose(f);
	return data;
}

complex float* cfl_load(const char* name, int D, long dims[])
{
	if (has_suffix(name, ".ra"))
		return load_ra(name, D, dims);

	char hdr[4096], bdy[4096];

	if ((int)sizeof(hdr) <= snprintf(hdr, sizeof(hdr), "%s.hdr", name)
	    || (int)sizeof(bdy) <= snprintf(bdy, sizeof(bdy), "%s.cfl", name))
		cli_error("Loading cfl file %s: name too long", name);

	FILE* f = fopen(hdr, "rb");

	if (NULL == f)
		cli_error("Loading cfl file %s", name);

	char text[16385];
	size_t r = fread(text, 1, sizeof(text) - 1, f);
	text[r] = '\0';
	fclose(f);

	parse_header(text, name, D, dims);

	long n = cfl_size(D, dims);
	complex float* data = malloc(sizeof(complex float) * (n > 0 ? n : 1));

	f = fopen(bdy, "rb");

	if (NULL == f)
		cli_error("Loading cfl file %s", name);

	read_exact(f, data, sizeof(complex float) * n, bdy);
	fclose(f);

	return data;
}

void cfl_write(const char* name, int D, const long dims[], const complex float* data)
{
	long n = cfl_size(D, dims);

	if (has_suffix(name, ".ra")) {

		FILE* f = fopen(name, "wb");

		if (NULL == f)
			cli_error("Creating ra file %s", name);

		uint64_t h[6] = { RA_MAGIC, 0, RA_TYPE_COMPLEX, 8, (uint64_t)n * 8, (uint64_t)D };
		write_exact(f, h, sizeof(h), name);

		for (int i = 0; i < D; i++) {
			uint64_t d = (uint64_t)dims[i];
			write_exact(f, &d, sizeof(d), name);
		}

		write_exact(f, data, sizeof(complex float) * n, name);
		fclose(f);
		return;
	}

	char hdr[4096], bdy[4096];

	if ((int)sizeof(hdr) <= snprintf(hdr, sizeof(hdr), "%s.hdr", name)
	    || (int)sizeof(bdy) <= snprintf(bdy, sizeof(bdy), "%s.cfl", name))
		cli_error("Creating cfl file %s: name too long", name);

	FILE* f = fopen(hdr, "w");

	if (NULL == f)
		cli_error("Creating cfl file %s", name);

	fprintf(f, "# Dimensions\n");

	for (int i = 0; i < D; i++)
		fprintf(f, "%ld ", dims[i]);

	fprintf(f, "\n");

	if (NULL != command_line)
		fprintf(f, "# Command\n%s\n", command_line);

	fprintf(f, "# Creator\nbart_amd\n");
	fclose(f);

	f = fopen(bdy, "wb");

	if (NULL == f)
		cli_error("Creating cfl file %s", name);

	write_exact(f, data, sizeof(complex float) * n, bdy);
	fclose(f);
}
