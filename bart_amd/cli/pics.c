/*
 * pics.c -- "bart pics" for non-Cartesian CG-SENSE with l2 regularisation
 * on the gfx950 library (src/pics.c:60-795, the -t path with CG).
 *
 *   pics -t traj [-l2] [-r lambda] [-i iter] [-p pattern] [-W img] [-S]
 *        [-w scaling] [-d level] [-g] [--gpu-gridding] [-U,--lowmem]
 *        [--no-toeplitz] [--psf_export file] [--psf_import file]
 *        [--tol t] [-B basis] [-T truth] kspace sensitivities output
 *
 * Steps as src/pics.c: the sampling pattern (from -p, else non-zero samples
 * over coils, misc/mri2.c:50) multiplies the k-space and is the NUFFT's
 * weights; the data are divided by a scale estimated from |A^H y|
 * (sense/optcom.c:74-109) unless -w; then CG on (A^H A + lambda I) x = A^H y
 * (lsqr2 with an l2 penalty, grecon/italgo.c:84-101) from zero or the -W
 * image; -S multiplies the result by the scale.  --tol (not in BART, whose
 * CG runs a fixed iteration count) stops at |r| <= tol |A^H y|.  -B: a
 * subspace basis [1 1 1 1 1 T K]; the image then has K coefficients in
 * dim 6 (src/pics.c:292-333).
 * Other regularisers (-l1, -R, wavelets, TV, ...) and algorithms are rejected.
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
		"Usage: pics -t <traj> [-l2] [-r lambda] [-i iter] [-p <pattern>] [-W <img>] [-S] [-w scaling]\n"
		"            [-d level] [-g] [--gpu-gridding] [-U,--lowmem] [--no-toeplitz]\n"
		"            [--psf_export <file>] [--psf_import <file>] [--tol t] [-B <basis>] [-T <truth>]\n"
		"            <kspace> <sensitivities> <output>\n"
		"\n"
		"Parallel-imaging reconstruction (non-Cartesian CG-SENSE with l2 regularisation).\n");
}

int main_pics(int argc, char* argv[])
{
	const char* traj_file = NULL;
	const char* pat_file = NULL;
	const char* start_file = NULL;
	const char* truth_file = NULL;
	const char* psf_ofile = NULL;
	const char* psf_ifile = NULL;
	const char* basis_file = NULL;
	float lambda = -1.f;
	int maxiter = 30;
	float scaling = 0.f;
	bool scale_im = false;
	float tol = 0.f;
	struct nufft_conf_s conf = nufft_conf_defaults;

	enum { O_GPUGRID = 256, O_NOTOEP, O_PSFEX, O_PSFIM, O_TOL, O_LOWMEM, O_NUFFTCONF, O_UNSUP };

	static const struct option lopts[] = {
		{ "gpu-gridding", no_argument, NULL, O_GPUGRID },
		{ "lowmem", no_argument, NULL, O_LOWMEM },
		{ "no-toeplitz", no_argument, NULL, O_NOTOEP },
		{ "psf_export", required_argument, NULL, O_PSFEX },
		{ "psf_import", required_argument, NULL, O_PSFIM },
		{ "tol", required_argument, NULL, O_TOL },
		{ "precond", no_argument, NULL, O_UNSUP },
		{ "fista", no_argument, NULL, O_UNSUP },
		{ "ist", no_argument, NULL, O_UNSUP },
		{ "admm", no_argument, NULL, O_UNSUP },
		{ "pridu", no_argument, NULL, O_UNSUP },
		{ "wavelet", required_argument, NULL, O_UNSUP },
		{ "shared-img-dims", required_argument, NULL, O_UNSUP },
		{ "motion-field", required_argument, NULL, O_UNSUP },
		{ "nufft-conf", required_argument, NULL, O_NUFFTCONF },
		{ NULL, 0, NULL, 0 },
	};

	optind = 1;
	int c, li = -1;
	float cclambda = 0.f;

	// short options of src/pics.c:137-199
	const char* shortopts = "l:r:R:cs:i:t:nNgp:b:eHDFJT:W:d:O:o:u:C:q:f:Imaw:SL:KB:P:MUh";

	while (-1 != (c = getopt_long(argc, argv, shortopts, lopts, &li))) {

		switch (c) {
		case 'l':
			if (0 != strcmp(optarg, "2"))
				unsupported_option("pics", (0 == strcmp(optarg, "1")) ? "-l1" : "-l");
			break;
		case 'r': lambda = parse_float(optarg, "-r"); break;
		case 'q': cclambda = parse_float(optarg, "-q"); break;
		case 'i': maxiter = parse_int(optarg, "-i"); break;
		case 't': traj_file = optarg; break;
		case 'g': break;
		case O_GPUGRID: break;
		case 'p': pat_file = optarg; break;
		case 'W': start_file = optarg; break;
		case 'd': set_debug_level(parse_int(optarg, "-d")); break;
		case 'w': scaling = parse_float(optarg, "-w"); break;
		case 'S': scale_im = true; break;
		case 'U':
		case O_LOWMEM: conf.lowmem = true; break;
		case O_NOTOEP: conf.toeplitz = false; break;
		case O_PSFEX: psf_ofile = optarg; break;
		case O_PSFIM: psf_ifile = optarg; break;
		case O_TOL: tol = parse_float(optarg, "--tol"); break;
		case O_NUFFTCONF: parse_nufft_conf(optarg, &conf); break;

		// options that do not change a CG reconstruction (step size,
		// wavelet/LLR/ADMM/IST settings, truth image for monitoring)
		case 's': case 'n': case 'N': case 'b': case 'e': case 'H':
		case 'D': case 'F': case 'J': case 'u': case 'C': case 'o':
			break;
		case 'T': truth_file = optarg; break;
		case 'B': basis_file = optarg; break;

		case O_UNSUP: unsupported_option("pics", lopts[li].name); break;
		case 'h': usage(); return 0;
		case '?': usage(); return 1;
		default: {
			char o[3] = { '-', (char)c, 0 };
			unsupported_option("pics", o);
		}
		}

		li = -1;
	}

	if (3 != argc - optind) {
		usage();
		cli_error("pics: expected 3 file arguments");
	}

	if (NULL == traj_file)
		cli_error("pics: only non-Cartesian data (-t) is supported");

	if (maxiter < 0)
		cli_error("pics: -i needs a non-negative iteration count");

	double start_time = timestamp();

	const char* ksp_file = argv[optind];
	const char* sens_file = argv[optind + 1];
	const char* out_file = argv[optind + 2];

	if (lambda < 0.f)
		lambda = 0.f;

	long ksp_dims[DIMS];
	complex float* kspace = cfl_load(ksp_file, DIMS, ksp_dims);

	long map_dims[DIMS];
	complex float* maps = cfl_load(sens_file, DIMS, map_dims);

	long traj_dims[DIMS];
	complex float* traj = cfl_load(traj_file, DIMS, traj_dims);

	// subspace basis [1 1 1 1 1 T K] (src/pics.c:292-333): the image holds
	// K coefficients in dim 6 instead of the T frames of dim 5
	long basis_dims[DIMS];
	complex float* basis = NULL;

	if (NULL != basis_file) {

		basis = cfl_load(basis_file, DIMS, basis_dims);

		for (int i = 0; i < DIMS; i++)
			if (TE_DIM != i && COEFF_DIM != i && 1 != basis_dims[i])
				cli_error("pics: the basis may vary over dims 5 and 6 only");
	}

	long max_dims[DIMS];
	memcpy(max_dims, ksp_dims, sizeof(max_dims));
	memcpy(max_dims, map_dims, 5 * sizeof(long));

	if (NULL != basis) {

		if (1 != ksp_dims[COEFF_DIM])
			cli_error("pics: k-space must have size 1 in dim 6 with a basis");

		if (basis_dims[TE_DIM] != ksp_dims[TE_DIM])
			cli_error("pics: basis and k-space do not match in dim 5");

		max_dims[COEFF_DIM] = basis_dims[COEFF_DIM];
		max_dims[TE_DIM] = 1;
	}

	long img_dims[DIMS];
	memcpy(img_dims, max_dims, sizeof(img_dims));
	img_dims[COIL_DIM] = 1;

	for (int i = 3; i < DIMS; i++)
		if (i != COIL_DIM && img_dims[i] != map_dims[i] && 1 != map_dims[i] && 1 != img_dims[i])
			cli_error("Dimensions of image and sensitivities do not match!\n");

	for (int i = 0; i < DIMS; i++)
		if (ksp_dims[i] != traj_dims[i] && 1 != ksp_dims[i] && 1 != traj_dims[i])
			cli_error("Dimensions of data and trajectory do not match!\n");

	if (1 != ksp_dims[MAPS_DIM])
		cli_error("pics: k-space must have size 1 in the maps dim");

	if (map_dims[MAPS_DIM] > 1)
		debug_printf(2, "%ld maps.\nESPIRiT reconstruction.\n", map_dims[MAPS_DIM]);

	debug_printf(2, "GPU reconstruction\n");

	// sampling pattern (src/pics.c:400-426)
	long pat_dims[DIMS];
	complex float* pattern = NULL;

	if (NULL != pat_file) {

		pattern = cfl_load(pat_file, DIMS, pat_dims);

		for (int i = 0; i < DIMS; i++)
			if (pat_dims[i] != ksp_dims[i] && 1 != pat_dims[i])
				cli_error("pics: pattern and k-space do not match in dim %d", i);

	} else {

		memcpy(pat_dims, ksp_dims, sizeof(pat_dims));
		pat_dims[COIL_DIM] = 1;
		pattern = malloc(sizeof(complex float) * (size_t)cfl_size(DIMS, pat_dims));
		estimate_pattern(ksp_dims, pattern, kspace);
	}

	{
		// kspace *= pattern (broadcast over the dims where the pattern is 1)
		long n = cfl_size(DIMS, ksp_dims);
		long pstr[DIMS], ps = 1;

		for (int i = 0; i < DIMS; i++) {
			pstr[i] = (1 == pat_dims[i]) ? 0 : ps;
			ps *= pat_dims[i];
		}

		for (long j = 0; j < n; j++) {

			long r = j, off = 0;

			for (int i = 0; i < DIMS; i++) {
				off += (r % ksp_dims[i]) * pstr[i];
				r /= ksp_dims[i];
			}

			kspace[j] *= pattern[off];
		}
	}

	if (NULL != psf_ifile && NULL == psf_ofile)
		conf.nopsf = true;

	const struct linop_s* nufft_op = NULL;
	const struct linop_s* forward_op = sense_nc_init(max_dims, map_dims, maps, ksp_dims, traj_dims, traj, &conf,
							 pat_dims, pattern, basis ? basis_dims : NULL, basis,
							 &nufft_op, 0);

	if (NULL != psf_ofile) {

		int D = nufft_get_psf_dims(nufft_op, 0, NULL);
		long psf_dims[DIMS + 1];
		nufft_get_psf_dims(nufft_op, D, psf_dims);
		complex float* psf = malloc(sizeof(complex float) * (size_t)cfl_size(D, psf_dims));
		nufft_get_psf(nufft_op, D, psf_dims, psf);
		cfl_write(psf_ofile, D, psf_dims, psf);
		free(psf);
	}

	if (NULL != psf_ifile) {

		long psf_dims[DIMS + 1];
		complex float* psf = cfl_load(psf_ifile, DIMS + 1, psf_dims);
		nufft_update_psf(nufft_op, DIMS + 1, psf_dims, psf);
		free(psf);
	}

	linop_free(nufft_op);

	long isize = cfl_size(DIMS, img_dims);
	long ksize = cfl_size(DIMS, ksp_dims);

	// scaling (src/pics.c:501-538)
	if (0.f == scaling) {

		complex float* adj = malloc(sizeof(complex float) * (size_t)isize);
		linop_adjoint(forward_op, DIMS, img_dims, adj, DIMS, ksp_dims, kspace);
		scaling = estimate_scaling_norm(isize, adj);
		free(adj);
	}

	if (0.f == scaling) {

		debug_printf(1, "Estimated scale is zero. Set to one.");
		scaling = 1.f;

	} else {

		debug_printf(3, "Inverse scaling of the data: %f\n", scaling);

		const complex float s = (complex float)(1. / scaling);

		for (long j = 0; j < ksize; j++)
			kspace[j] *= s;
	}

	complex float* image = calloc((size_t)isize, sizeof(complex float));
	bool warm = false;

	if (NULL != start_file) {

		long sd[DIMS];
		complex float* st = cfl_load(start_file, DIMS, sd);

		for (int i = 0; i < DIMS; i++)
			if (sd[i] != img_dims[i])
				cli_error("pics: warm start image has the wrong dimensions");

		debug_printf(3, "Warm start: %s\n", start_file);

		// with -S the warm start is taken to be rescaled too (src/pics.c:616-618)
		const complex float s = (complex float)(1. / scaling);

		for (long j = 0; j < isize; j++)
			image[j] = (scale_im && 0.f != scaling) ? st[j] * s : st[j];

		free(st);
		warm = true;
	}

	debug_printf(2, "Regularization terms: %d, Supporting variables: 0\n", (lambda > 0.f) ? 1 : 0);
	debug_printf(2, "conjugate gradients\n");

	complex float* b = malloc(sizeof(complex float) * (size_t)isize);
	linop_adjoint(forward_op, DIMS, img_dims, b, DIMS, ksp_dims, kspace);

	float* trace = calloc((size_t)(maxiter > 0 ? maxiter : 1), sizeof(float));
	float* errs = NULL;

	if (NULL != truth_file) {

		// the monitor compares the iterates with the truth image as given,
		// without the data scaling (src/pics.c:548-575)
		long td[DIMS];
		complex float* truth = cfl_load(truth_file, DIMS, td);

		for (int i = 0; i < DIMS; i++)
			if (td[i] != img_dims[i])
				cli_error("pics: truth image has the wrong dimensions");

		errs = calloc((size_t)(maxiter > 0 ? maxiter : 1), sizeof(float));
		bart_amd_cg_monitor(forward_op, truth, errs);
		free(truth);
	}

	struct cg_stats_s st;
	// lsqr's lambda (cclambda) and the l2 penalty both add a multiple of the
	// identity to the normal operator (iter/lsqr.c:52-60, iter/italgos.c:497)
	lsqr_conjgrad(forward_op, maxiter, cclambda + lambda, tol, DIMS, img_dims, image, b, warm, false, trace, &st);

	for (int i = 0; i < st.iterations; i++)
		debug_printf(5, "#CG %3d: |r| = %e\n", i + 1, trace[i]);

	// iter/monitor.c:68 (no objective in CG: -1)
	if (NULL != errs)
		for (int i = 0; i < st.iterations; i++)
			debug_printf(2, "[Iter %04d] Objective: %f, Error: %f\n", i + 1, -1., (double)errs[i]);

	debug_printf(3, "CG: %d iterations, |r| / |b| = %e, %.3f ms on the GPU\n", st.iterations,
		     st.bnorm > 0.f ? st.resid / st.bnorm : 0.f, st.time_ms);

	if (scale_im) {

		const complex float s = scaling;

		for (long j = 0; j < isize; j++)
			image[j] *= s;
	}

	cfl_write(out_file, DIMS, img_dims, image);

	linop_free(forward_op);

	free(trace);
	free(errs);
	free(b);
	free(image);
	free(pattern);
	free(basis);
	free(traj);
	free(maps);
	free(kspace);

	debug_printf(2, "Total Time: %f\n", timestamp() - start_time);

	return 0;
}
