/*
 * ecalib.c -- "bart ecalib" on the gfx950 library (src/ecalib.c:47-263):
 * ESPIRiT coil sensitivities from Cartesian k-space.
 *
 *   ecalib [-t threshold] [-c crop] [-k kx:ky:kz] [-r cx:cy:cz] [-m maps]
 *          [-P] [-i orthiter] [-g] [-d level] kspace sensitivities [ev-maps]
 *
 * The calibration region is extracted as extract_calib does, the kernel
 * size is 1 along dims of size 1, and calib runs on the GPU.  Soft
 * weighting (-W), automatic thresholds (-a, -v), intensity and phase
 * normalisation (-I, -N), soft crop (-S), Nystroem (-A), the
 * eigen-decomposition without orthogonal iteration (-O), only the first
 * part (-1) and the fixed-centre region (-C) are rejected with an error.
 */
#include "tools.h"
#include "cfl.h"
#include "bart_amd.h"

#include <complex.h>
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void usage(void)
{
	fprintf(stderr,
		"Usage: ecalib [-t threshold] [-c crop_value] [-k ksize] [-r cal_size] [-m maps] [-P]\n"
		"              <kspace> <sensitivities> [<ev-maps>]\n"
		"\n"
		"Estimate coil sensitivities using ESPIRiT calibration.\n");
}

int main_ecalib(int argc, char* argv[])
{
	struct ecalib_conf conf = ecalib_defaults;
	long calsize[3] = { 24, 24, 24 };
	int maps = 2;
	bool print_svals = false;

	optind = 1;
	int c;

	while (-1 != (c = getopt(argc, argv, "t:c:k:K:r:R:m:SWI1PNAOi:b:VCgp:n:v:ae:d:h"))) {

		switch (c) {
		case 't': conf.threshold = parse_float(optarg, "-t"); break;
		case 'c': conf.crop = parse_float(optarg, "-c"); break;
		case 'k':
		case 'K': parse_vec3(optarg, conf.kdims, "-k"); break;
		case 'r':
		case 'R': parse_vec3(optarg, calsize, "-r"); break;
		case 'm': maps = parse_int(optarg, "-m"); break;
		case 'P': conf.rotphase = false; break;
		case 'i': conf.num_orthiter = parse_int(optarg, "-i"); break;
		case 'V': print_svals = true; break;
		case 'g': break;
		case 'e': conf.econdim = parse_int(optarg, "-e"); break;
		case 'd': set_debug_level(parse_int(optarg, "-d")); break;
		case 'h': usage(); return 0;
		case 'S': case 'W': case 'I': case '1': case 'N': case 'A': case 'O': case 'b': case 'C':
		case 'p': case 'n': case 'v': case 'a': {
			char o[3] = { '-', (char)c, 0 };
			unsupported_option("ecalib", o);
			break;
		}
		default: usage(); return 1;
		}
	}

	if (argc - optind < 2 || argc - optind > 3) {
		usage();
		cli_error("ecalib: expected 2 or 3 file arguments");
	}

	if (maps < 1)
		cli_error("ecalib: -m needs at least one map");

	long ksp_dims[DIMS];
	complex float* ksp = cfl_load(argv[optind], DIMS, ksp_dims);

	if (1 != ksp_dims[MAPS_DIM])
		cli_error("MAPS dimension is not of size one.");

	if (maps > ksp_dims[COIL_DIM])
		cli_error("ecalib: more maps than coils");

	long cal_dims[DIMS];
	complex float* cal = extract_calib(cal_dims, calsize, ksp_dims, ksp, false);

	if (NULL == cal)
		cli_error("ecalib: %s", bart_amd_last_error() ? bart_amd_last_error() : "calibration failed");

	for (int i = 0; i < 3; i++)
		if (1 == ksp_dims[i])
			conf.kdims[i] = 1;

	for (int i = 0; i < 3; i++)
		if ((1 == cal_dims[i]) && (1 != ksp_dims[i]))
			cli_error("Calibration region not found!");

	debug_printf(2, "Calibration region...  (size: %ldx%ldx%ld)\n", cal_dims[0], cal_dims[1], cal_dims[2]);

	long out_dims[DIMS];
	long map_dims[DIMS];

	for (int i = 0; i < DIMS; i++) {
		out_dims[i] = 1;
		map_dims[i] = 1;
		if (i < 3 && 1 < conf.kdims[i]) {
			out_dims[i] = ksp_dims[i];
			map_dims[i] = ksp_dims[i];
		}
	}

	out_dims[COIL_DIM] = ksp_dims[COIL_DIM];
	out_dims[MAPS_DIM] = maps;
	map_dims[MAPS_DIM] = maps;

	int K = (int)(conf.kdims[0] * conf.kdims[1] * conf.kdims[2] * ksp_dims[COIL_DIM]);
	float* svals = calloc((size_t)K, sizeof(float));
	complex float* out = calloc((size_t)cfl_size(DIMS, out_dims), sizeof(complex float));
	complex float* emaps = calloc((size_t)cfl_size(DIMS, map_dims), sizeof(complex float));

	calib(&conf, out_dims, out, emaps, K, svals, cal_dims, cal);

	cfl_write(argv[optind + 1], DIMS, out_dims, out);

	if (argc - optind == 3)
		cfl_write(argv[optind + 2], DIMS, map_dims, emaps);

	if (print_svals)
		for (int i = 0; i < K; i++)
			printf("SVALS %d %f\n", i, svals[i]);

	free(svals);
	free(out);
	free(emaps);
	free(cal);
	free(ksp);

	return 0;
}
