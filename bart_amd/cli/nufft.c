/*
 * nufft.c -- "bart nufft" on the gfx950 library (src/nufft.c:44-330).
 *
 *   nufft [-a | -i] [-x x:y:z] [-t | -r] [-l lambda] [-m iter] [-P] [-g] [-1]
 *         [-p pattern] [-o 2] [-w width] [--lowmem] [--zero-mem]
 *         [--no-precomp] traj input output
 *
 * forward: image [X Y Z C ...] -> k-space [1 R S C ...]
 * -a: adjoint; -i: inverse (CG on (A^H A + lambda I) x = A^H y, Toeplitz
 * normal operator unless -r, 50 iterations by default as
 * iter_conjgrad_defaults, iter/iter.c:51-60).
 * -s: the direct NUDFT instead of the NUFFT (noncart/nudft.c).
 * -c: Strang's circulant preconditioner for -i (noncart/precond.c).
 * -B: a subspace basis [1 1 1 1 1 T K] (src/nufft.c:169-176, 208-212,
 * 284-288): coil images [X Y Z C 1 1 K], k-space [1 R S C 1 T].
 * -1: the non-decomposed NUFFT whose image is the oversampled grid itself
 * (an estimated image size is doubled, src/nufft.c:197-202); it has no
 * Toeplitz operator (-i needs -r).  --nufft-conf: BART's sub-options.  -g
 * is accepted (every path runs on the GPU).  -F/-T (B0) are rejected with
 * an error.
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
		"Usage: nufft [-a,-i] [-x x:y:z] [-t,-r] [-l lambda] [-m iter] [-P] [-s] [-g] [-1] [-p <pattern>]\n"
		"             [-o os] [-w width] [-c] [-B <basis>] [--lowmem] [--zero-mem] [--no-precomp]\n"
		"             <traj> <input> <output>\n"
		"\n"
		"Perform non-uniform Fast Fourier Transform.\n");
}

int main_nufft(int argc, char* argv[])
{
	bool adjoint = false;
	bool inverse = false;
	bool precond = false;
	bool dft = false;
	bool precomp = true;
	long coilim_vec[3] = { 0, 0, 0 };
	float lambda = 0.f;
	int maxiter = 50;
	const char* pattern_file = NULL;
	const char* basis_file = NULL;
	struct nufft_conf_s conf = nufft_conf_defaults;

	enum { O_LOWMEM = 256, O_ZEROMEM, O_NOPRECOMP, O_NUFFTCONF };

	static const struct option lopts[] = {
		{ "lowmem", no_argument, NULL, O_LOWMEM },
		{ "zero-mem", no_argument, NULL, O_ZEROMEM },
		{ "no-precomp", no_argument, NULL, O_NOPRECOMP },
		{ "oversampling", required_argument, NULL, 'o' },
		{ "width", required_argument, NULL, 'w' },
		{ "nufft-conf", required_argument, NULL, O_NUFFTCONF },
		{ NULL, 0, NULL, 0 },
	};

	optind = 1;
	int c;

	while (-1 != (c = getopt_long(argc, argv, "aix:d:D:trcl:m:Psg1B:p:o:w:F:T:h", lopts, NULL))) {

		switch (c) {
		case 'a': adjoint = true; break;
		case 'i': inverse = true; break;
		case 'x':
		case 'd':
		case 'D': parse_vec3(optarg, coilim_vec, "-x"); break;
		case 't': conf.toeplitz = true; break;
		case 'r': conf.toeplitz = false; break;
		case 'l': lambda = parse_float(optarg, "-l"); break;
		case 'm': maxiter = parse_int(optarg, "-m"); break;
		case 'P': conf.periodic = true; break;
		case 'g': break;
		case 'p': pattern_file = optarg; break;
		case 'o': conf.os = parse_float(optarg, "-o"); break;
		case 'w': conf.width = parse_float(optarg, "-w"); break;
		case O_LOWMEM: conf.lowmem = true; break;
		case O_ZEROMEM: conf.zero_overhead = true; break;
		case O_NOPRECOMP: precomp = false; break;
		case 'c': precond = true; break;
		case 's': dft = true; break;
		case '1': conf.decomp = false; break;
		case 'B': basis_file = optarg; break;
		case 'F': unsupported_option("nufft", "-F"); break;
		case 'T': unsupported_option("nufft", "-T"); break;
		case O_NUFFTCONF: parse_nufft_conf(optarg, &conf); break;
		case 'h': usage(); return 0;
		default: usage(); return 1;
		}
	}

	if (3 != argc - optind) {
		usage();
		cli_error("nufft: expected 3 file arguments");
	}

	const char* traj_file = argv[optind];
	const char* in_file = argv[optind + 1];
	const char* out_file = argv[optind + 2];

	if (maxiter < 0)
		cli_error("nufft: -m needs a non-negative iteration count");

	if (adjoint && inverse)
		cli_error("Adjoint and inverse requested at the same time.\n");

	if (!precomp) {
		conf.lowmem = true;
		conf.precomp_linphase = false;
		conf.precomp_fftmod = false;
		conf.precomp_roll = false;
	}

	// avoid computing the PSF if not necessary (src/nufft.c:134-136)
	if (!inverse)
		conf.toeplitz = false;

	long traj_dims[DIMS];
	complex float* traj = cfl_load(traj_file, DIMS, traj_dims);

	if (3 != traj_dims[0])
		cli_error("nufft: trajectory must have 3 coordinates in dim 0");

	long coilest[3];
	estimate_im_dims(traj_dims, traj, coilest);

	if (8 >= coilest[0] * coilest[1] * coilest[2])
		debug_printf(1, "\tThe estimated image size %ldx%ldx%ld is very small.\n"
				"\tDid you scale your trajectory correctly?\n"
				"\tThe unit of measurement is pixel_size / FOV.\n",
			     coilest[0], coilest[1], coilest[2]);

	long pattern_dims[DIMS];
	complex float* pattern = NULL;

	if (NULL != pattern_file)
		pattern = cfl_load(pattern_file, DIMS, pattern_dims);

	if (dft && NULL != pattern)
		cli_error("nufft: -s does not take a pattern (-p)");

	long basis_dims[DIMS];
	complex float* basis = NULL;

	if (NULL != basis_file) {

		basis = cfl_load(basis_file, DIMS, basis_dims);

		for (int i = 0; i < DIMS; i++)
			if (TE_DIM != i && COEFF_DIM != i && 1 != basis_dims[i])
				cli_error("nufft: the basis may vary over dims 5 and 6 only");

		if (dft)
			cli_error("nufft: -s does not take a basis (-B)");

		if (precond)
			cli_error("nufft: -c does not take a basis (-B)");
	}

	if (inverse || adjoint) {

		long ksp_dims[DIMS];
		complex float* ksp = cfl_load(in_file, DIMS, ksp_dims);

		if (1 != ksp_dims[0])
			cli_error("nufft: k-space must have size 1 in dim 0");

		for (int i = 1; i < 3; i++)
			if (ksp_dims[i] != traj_dims[i] && 1 != ksp_dims[i] && 1 != traj_dims[i])
				cli_error("nufft: k-space and trajectory do not match in dim %d", i);

		long coilim_dims[DIMS];

		for (int i = 0; i < DIMS; i++)
			coilim_dims[i] = 1;

		if (0 == coilim_vec[0] * coilim_vec[1] * coilim_vec[2]) {

			for (int i = 0; i < 3; i++)
				coilim_dims[i] = (!conf.decomp && coilest[i] > 1) ? 2 * coilest[i] : coilest[i];

			debug_printf(2, "Est. image size: %ldx%ldx%ld\n", coilim_dims[0], coilim_dims[1], coilim_dims[2]);

		} else {

			for (int i = 0; i < 3; i++)
				coilim_dims[i] = coilim_vec[i];
		}

		for (int i = 3; i < DIMS; i++)
			coilim_dims[i] = ksp_dims[i];

		if (NULL != basis) {

			coilim_dims[COEFF_DIM] = basis_dims[COEFF_DIM];
			coilim_dims[TE_DIM] = 1;
		}

		long n = cfl_size(DIMS, coilim_dims);
		complex float* img = calloc((size_t)n, sizeof(complex float));

		const struct linop_s* op = dft ? nudft_create(DIMS, 7, ksp_dims, coilim_dims, traj_dims, traj, NULL, NULL,
							      NULL, NULL)
					       : nufft_create2(DIMS, ksp_dims, coilim_dims, traj_dims, traj,
							       pattern ? pattern_dims : NULL, pattern,
								       basis ? basis_dims : NULL, basis, conf);

		if (inverse) {

			complex float* b = calloc((size_t)n, sizeof(complex float));
			linop_adjoint(op, DIMS, coilim_dims, b, DIMS, ksp_dims, ksp);

			struct cg_stats_s st;

			// -c: Strang's circulant preconditioner with the Toeplitz
			// operator (src/nufft.c:250-263)
			if (precond && conf.toeplitz && !dft) {

				const struct operator_s* pre = nufft_precond_create(op);
				lsqr_conjgrad_precond(op, pre, maxiter, lambda, 0.f, DIMS, coilim_dims, img, b, false, NULL,
						      &st);
				operator_free(pre);

			} else {

				lsqr_conjgrad(op, maxiter, lambda, 0.f, DIMS, coilim_dims, img, b, false, false, NULL, &st);
			}

			debug_printf(3, "CG: %d iterations, |r| / |b| = %e, %.3f ms\n", st.iterations,
				     st.bnorm > 0.f ? st.resid / st.bnorm : 0.f, st.time_ms);
			free(b);

		} else {

			linop_adjoint(op, DIMS, coilim_dims, img, DIMS, ksp_dims, ksp);
		}

		linop_free(op);
		cfl_write(out_file, DIMS, coilim_dims, img);
		free(img);
		free(ksp);

	} else {

		long coilim_dims[DIMS];
		complex float* img = cfl_load(in_file, DIMS, coilim_dims);

		long ksp_dims[DIMS];

		for (int i = 0; i < DIMS; i++)
			ksp_dims[i] = (i < 3) ? ((0 == i) ? 1 : traj_dims[i]) : coilim_dims[i];

		if (NULL != basis) {

			ksp_dims[TE_DIM] = basis_dims[TE_DIM];
			ksp_dims[COEFF_DIM] = 1;
		}

		long m = cfl_size(DIMS, ksp_dims);
		complex float* ksp = calloc((size_t)m, sizeof(complex float));

		const struct linop_s* op = dft ? nudft_create(DIMS, 7, ksp_dims, coilim_dims, traj_dims, traj, NULL, NULL,
							      NULL, NULL)
					       : nufft_create2(DIMS, ksp_dims, coilim_dims, traj_dims, traj,
							       pattern ? pattern_dims : NULL, pattern,
								       basis ? basis_dims : NULL, basis, conf);

		linop_forward(op, DIMS, ksp_dims, ksp, DIMS, coilim_dims, img);
		linop_free(op);

		cfl_write(out_file, DIMS, ksp_dims, ksp);
		free(ksp);
		free(img);
	}

	free(traj);
	free(pattern);
	free(basis);

	return 0;
}
