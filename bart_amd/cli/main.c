/*
 * main.c -- the "bart" command of the gfx950 NUFFT / CG-SENSE library:
 * "bart <tool> [options] files..." (src/bart.c:45-250 dispatch).  Invoked
 * under a tool's name (a link named "nufft" or "pics") it runs that tool.
 */
#include "tools.h"
#include "cfl.h"
#include "bart_amd.h"

#include <libgen.h>
#include <stdio.h>
#include <string.h>

struct tool {
	const char* name;
	int (*main)(int argc, char* argv[]);
};

static const struct tool tools[] = {
	{ "ecalib", main_ecalib },
	{ "nufft", main_nufft },
	{ "pics", main_pics },
	{ "traj", main_traj },
};

static int version(void)
{
	printf("%s\n", bart_amd_version());
	return 0;
}

static void list(void)
{
	fprintf(stderr, "Usage: bart <command> ...\n\nAvailable commands:\n");
	for (size_t i = 0; i < sizeof(tools) / sizeof(tools[0]); i++)
		fprintf(stderr, "  %s\n", tools[i].name);
	fprintf(stderr, "  version\n");
}

static int run(const char* name, int argc, char* argv[])
{
	for (size_t i = 0; i < sizeof(tools) / sizeof(tools[0]); i++)
		if (0 == strcmp(name, tools[i].name))
			return tools[i].main(argc, argv);

	if (0 == strcmp(name, "version"))
		return version();

	list();
	cli_error("Unknown bart command: \"%s\".\n", name);
}

int main(int argc, char* argv[])
{
	cfl_set_command(argc, argv);

	char self[4096];
	snprintf(self, sizeof(self), "%s", argv[0]);
	const char* base = basename(self);

	if (0 != strcmp(base, "bart"))
		return run(base, argc, argv);

	if (argc < 2) {
		list();
		return 1;
	}

	return run(argv[1], argc - 1, argv + 1);
}
