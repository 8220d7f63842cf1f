/*
 * bart_oracle.c -- CPU restatement of BART's non-Cartesian NUFFT, Toeplitz
 * normal operator and CG-SENSE solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * bart_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU baseline
 * leg may load it.  Nothing in the product path links it.
 *
 * Layout (BART column-major, dimensions collapsed):
 *   image / coil image   cim[c][z][y][x]     (x fastest), N = X*Y*Z per coil
 *   k-space              ksp[c][m]           m = readout + R*spoke, M per coil
 *   trajectory           traj[m][3]          complex float, coordinate in the
 *                                            real part (cycles/FOV)
 *   sub-grids            grid[s][c][z][y][x] s = decomposition index
 *
 * Every function cites the reference file:line it restates.  Arithmetic
 * follows the reference's float/double choices where they matter for
 * rounding; tables that BART builds in FP32 from FP32 phases are built here in
 * FP64 and rounded once (a difference of a few ulp).
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef float complex cf;

#define KB_SIZE 1000

/* ------------------------------------------------------------------------ */
/* Kaiser-Bessel window                                                     */
/* ------------------------------------------------------------------------ */

/* Modified Bessel function of the first kind, order 0, by its power series
 * sum_k ((x/2)^k / k!)^2.  BART (num/specfun.c:120-129) uses a rational
 * approximation of the same function; both agree to ~1e-15 relative for
 * the arguments used here (|x| < 40). */
double orc_bessel_i0(double x)
{
	double q = 0.25 * x * x;
	double term = 1.;
	double sum = 1.;

	for (int k = 1; k < 500; k++) {

		term *= q / ((double)k * (double)k);
		sum += term;

		if (term < 1.e-17 * sum)
			break;
	}

	return sum;
}

/* noncart/grid.c:516-519 */
double orc_calc_beta(float os, float width)
{
	return M_PI * sqrt(pow((double)(width / os) * (os - 0.5), 2.) - 0.8);
}

/* noncart/grid.c:70-76 */
static double kb(double beta, double i0beta, double x)
{
	if (fabs(x) >= 0.5)
		return 0.;

	return orc_bessel_i0(beta * sqrt(1. - pow(2. * x, 2.))) / i0beta;
}

/* noncart/grid.c:86-97: table[i] = kb(beta, i / (n - 1) / 2), n = 1000 */
void orc_kb_table(double beta, float table[KB_SIZE + 1])
{
	double i0beta = orc_bessel_i0(beta);

	for (int i = 0; i < KB_SIZE + 1; i++)
		table[i] = (float)kb(beta, i0beta, (double)i / (double)(KB_SIZE - 1) / 2.);
}

/* noncart/grid.c:120-149 (intlookup, lerp) */
static inline float intlookup(const float table[KB_SIZE + 1], float x)
{
	x *= 2;

	int index = (int)(x * (KB_SIZE - 1));
	float fpart = x * (KB_SIZE - 1) - (float)index;

	return (float)((1. - fpart) * table[index] + fpart * table[index + 1]);
}

/* noncart/grid.c:100-117 */
static double ftkb(double beta, double i0beta, double x)
{
	double a = pow(beta, 2.) - pow(M_PI * x, 2.);

	if (0. == a)
		return 1. / i0beta;

	if (a > 0)
		return (sinh(sqrt(a)) / sqrt(a)) / i0beta;
	else
		return (sin(sqrt(-a)) / sqrt(-a)) / i0beta;
}

static double rolloff(double x, double beta, double i0beta, double width)
{
	return 1. / ftkb(beta, i0beta, x * width) / width;
}

/* noncart/grid.c:522-531 */
static float rpos(int d, int i, float os)
{
	if (1 == d)
		return 0.;

	int od = os * d;
	int oi = i + (od / 2 - d / 2);

	return ((float)oi - (float)(od / 2)) / (float)od;
}

/* noncart/grid.c:537-556: deapodisation over an image of dims[3] */
void orc_rolloff(float os, float width, double beta, const long dims[3], cf* dst)
{
	double i0beta = orc_bessel_i0(beta);

	#pragma omp parallel for collapse(2) schedule(static)
	for (long z = 0; z < dims[2]; z++)
		for (long y = 0; y < dims[1]; y++)
			for (long x = 0; x < dims[0]; x++)
				dst[x + dims[0] * (y + z * dims[1])]
					= (float)((dims[0] > 1 ? rolloff(rpos(dims[0], x, os), beta, i0beta, width) : 1.)
					* (dims[1] > 1 ? rolloff(rpos(dims[1], y, os), beta, i0beta, width) : 1.)
					* (dims[2] > 1 ? rolloff(rpos(dims[2], z, os), beta, i0beta, width) : 1.));
}


/* ------------------------------------------------------------------------ */
/* Gridding                                                                 */
/* ------------------------------------------------------------------------ */

struct orc_grid_conf {

	float os;
	float width;
	bool periodic;
	float shift[3];
};

/* Tap range and weights of one sample along one dimension:
 * noncart/grid.c:415-441 (range, clamp / wrap) and :453-457 (weights). */
struct taps1d {

	int n;
	long idx[16];
	float w[16];
};

static bool taps_1d(struct taps1d* t, float pos, long dim, bool periodic, float width, const float* table)
{
	t->n = 0;

	if (1 == dim) {

		t->n = 1;
		t->idx[0] = 0;
		t->w[0] = intlookup(table, fabsf(0.f - pos) / width);
		return true;
	}

	int sti = (int)ceil(pos - 0.5 * width);
	int eni = (int)floor(pos + 0.5 * width);
	int off = 0;

	if (sti > eni)
		return false;

	if (!periodic) {

		if (sti < 0)
			sti = 0;
		if (eni > dim - 1)
			eni = dim - 1;

	} else {

		while (sti + off < 0)
			off += dim;
	}

	for (int w = sti; w <= eni; w++) {

		float frac = fabsf((float)w - pos);
		t->w[t->n] = intlookup(table, frac / width);
		t->idx[t->n] = (w + off) % dim;
		t->n++;
	}

	return true;
}

static void sample_pos(const struct orc_grid_conf* conf, const long gdims[3], const cf* tr, float pos[3])
{
	for (int j = 0; j < 3; j++) {

		/* creal() of grid.c:181-183 widens to double: the sum and the
		 * product are double, rounded to float once */
		pos[j] = (float)(conf->os * ((double)crealf(tr[j]) + conf->shift[j]));
		pos[j] += (gdims[j] > 1) ? ((float)(gdims[j] / 2)) : 0.f;
	}
}

/* Adjoint gridding ("grid" in BART, k-space -> grid, scatter):
 * noncart/grid.c:202-253 and grid_point :469-490.  The loop nest is
 * z (outer), y, x (inner) with the weight accumulated in that order
 * (grid_point_gen recursion :443-464).  Zero samples (all coils) are
 * skipped as in :241-250.  grid is [C][gdims] and accumulated into.
 *
 * BART runs the sample loop in parallel and adds with atomics
 * (grid.c:482-485), so the order of the additions into one grid point, and
 * with it the rounding, varies from run to run.  The oracle fixes the order
 * to that of the serial loop: the grid is cut into slabs along its last
 * dimension of size > 1, each slab gets the list of the samples whose taps
 * reach it (in sample order), and a thread adds a slab's samples, taps
 * inside the slab only.  Every grid point then receives its contributions
 * in sample order, whatever the thread count. */
static void grid_core(const struct orc_grid_conf* conf, const float* table, long M, const cf* traj,
		const long gdims[3], int C, cf* grid, double complex* grid64, long gstride_c,
		const cf* ksp, long kstride_c)
{
	int D = 2;

	while (D > 0 && 1 == gdims[D])
		D--;

	long nslab = (gdims[D] < 256) ? gdims[D] : 256;
	int* slab_of = malloc(sizeof(int) * (size_t)gdims[D]);

	for (long sl = 0; sl < nslab; sl++)
		for (long r = sl * gdims[D] / nslab; r < (sl + 1) * gdims[D] / nslab; r++)
			slab_of[r] = (int)sl;

	/* samples per slab, then the lists in sample order */
	long* cnt = calloc((size_t)nslab + 1, sizeof(long));
	uint8_t* nsl = malloc((size_t)M);
	int (*sls)[16] = malloc(sizeof(int[16]) * (size_t)M);

	#pragma omp parallel for schedule(static)
	for (long m = 0; m < M; m++) {

		nsl[m] = 0;

		bool skip = true;

		for (int c = 0; c < C; c++)
			skip = skip && (0. == ksp[m + c * kstride_c]);

		if (skip)
			continue;

		float pos[3];
		sample_pos(conf, gdims, traj + 3 * m, pos);

		struct taps1d t;

		if (!taps_1d(&t, pos[D], gdims[D], conf->periodic, conf->width, table))
			continue;

		for (int i = 0; i < t.n; i++) {

			int sl = slab_of[t.idx[i]];
			bool seen = false;

			for (int j = 0; j < nsl[m]; j++)
				seen = seen || (sls[m][j] == sl);

			if (!seen)
				sls[m][nsl[m]++] = sl;
		}
	}

	for (long m = 0; m < M; m++)
		for (int j = 0; j < nsl[m]; j++)
			cnt[sls[m][j] + 1]++;

	for (long sl = 0; sl < nslab; sl++)
		cnt[sl + 1] += cnt[sl];

	long* fill = malloc(sizeof(long) * (size_t)nslab);
	memcpy(fill, cnt, sizeof(long) * (size_t)nslab);
	long* list = malloc(sizeof(long) * (size_t)(cnt[nslab] > 0 ? cnt[nslab] : 1));

	for (long m = 0; m < M; m++)
		for (int j = 0; j < nsl[m]; j++)
			list[fill[sls[m][j]]++] = m;

	#pragma omp parallel for schedule(dynamic, 1)
	for (long sl = 0; sl < nslab; sl++) {

		long lo = sl * gdims[D] / nslab;
		long hi = (sl + 1) * gdims[D] / nslab;

		for (long e = cnt[sl]; e < cnt[sl + 1]; e++) {

			long m = list[e];

			float pos[3];
			sample_pos(conf, gdims, traj + 3 * m, pos);

			struct taps1d t[3];

			if (!taps_1d(&t[0], pos[0], gdims[0], conf->periodic, conf->width, table)
			    || !taps_1d(&t[1], pos[1], gdims[1], conf->periodic, conf->width, table)
			    || !taps_1d(&t[2], pos[2], gdims[2], conf->periodic, conf->width, table))
				continue;

			for (int iz = 0; iz < t[2].n; iz++) {

				if (2 == D && (t[2].idx[iz] < lo || t[2].idx[iz] >= hi))
					continue;

				float dz = 1.f * t[2].w[iz];

				for (int iy = 0; iy < t[1].n; iy++) {

					if (1 == D && (t[1].idx[iy] < lo || t[1].idx[iy] >= hi))
						continue;

					float dy = dz * t[1].w[iy];

					for (int ix = 0; ix < t[0].n; ix++) {

						if (0 == D && (t[0].idx[ix] < lo || t[0].idx[ix] >= hi))
							continue;

						float d = dy * t[0].w[ix];
						long ind = t[0].idx[ix] + gdims[0] * (t[1].idx[iy] + gdims[1] * t[2].idx[iz]);

						if (NULL != grid64) {

							/* products and sums in double: the
							 * reference without FP32 rounding */
							double d64 = (double)t[2].w[iz] * (double)t[1].w[iy] * (double)t[0].w[ix];

							for (int c = 0; c < C; c++) {

								cf v = ksp[m + c * kstride_c];
								double* dst = (double*)&grid64[ind + c * gstride_c];

								dst[0] += (double)crealf(v) * d64;
								dst[1] += (double)cimagf(v) * d64;
							}

							continue;
						}

						for (int c = 0; c < C; c++) {

							cf v = ksp[m + c * kstride_c];
							float* dst = (float*)&grid[ind + c * gstride_c];

							dst[0] += crealf(v) * d;
							dst[1] += cimagf(v) * d;
						}
					}
				}
			}
		}
	}

	free(list);
	free(fill);
	free(sls);
	free(nsl);
	free(cnt);
	free(slab_of);
}

void orc_grid(const struct orc_grid_conf* conf, const float* table, long M, const cf* traj,
		const long gdims[3], int C, cf* grid, long gstride_c, const cf* ksp, long kstride_c)
{
	grid_core(conf, table, M, traj, gdims, C, grid, NULL, gstride_c, ksp, kstride_c);
}

/* Forward gridding ("gridH" in BART, grid -> k-space, gather):
 * noncart/grid.c:153-199 and grid_pointH :494-512.  ksp is accumulated into. */
void orc_gridH(const struct orc_grid_conf* conf, const float* table, long M, const cf* traj,
		int C, cf* ksp, long kstride_c, const long gdims[3], const cf* grid, long gstride_c)
{
	#pragma omp parallel for schedule(static)
	for (long m = 0; m < M; m++) {

		float pos[3];
		sample_pos(conf, gdims, traj + 3 * m, pos);

		struct taps1d t[3];

		if (!taps_1d(&t[0], pos[0], gdims[0], conf->periodic, conf->width, table)
		    || !taps_1d(&t[1], pos[1], gdims[1], conf->periodic, conf->width, table)
		    || !taps_1d(&t[2], pos[2], gdims[2], conf->periodic, conf->width, table))
			continue;

		cf val[C];

		for (int c = 0; c < C; c++)
			val[c] = 0.;

		for (int iz = 0; iz < t[2].n; iz++) {

			float dz = 1.f * t[2].w[iz];

			for (int iy = 0; iy < t[1].n; iy++) {

				float dy = dz * t[1].w[iy];

				for (int ix = 0; ix < t[0].n; ix++) {

					float d = dy * t[0].w[ix];
					long ind = t[0].idx[ix] + gdims[0] * (t[1].idx[iy] + gdims[1] * t[2].idx[iz]);

					for (int c = 0; c < C; c++) {

						cf g = grid[ind + c * gstride_c];
						val[c] = CMPLXF(crealf(val[c]) + crealf(g) * d, cimagf(val[c]) + cimagf(g) * d);
					}
				}
			}
		}

		for (int c = 0; c < C; c++)
			ksp[m + c * kstride_c] += val[c];
	}
}


/* ------------------------------------------------------------------------ */
/* FFT (the role FFTW plays in BART: num/fft_plan.c:188-281)                */
/* Unnormalised complex FFT, sign -1 forward, +1 backward, over every       */
/* dimension of size > 1.  Mixed-radix Stockham in FP32 with FP64 twiddles. */
/* ------------------------------------------------------------------------ */

struct fft1d_plan {

	int n;
	int nr;
	int radix[40];
	cf* tw;		/* exp(-2 pi i k / n), k < n */
};

static void fft1d_plan_init(struct fft1d_plan* p, int n)
{
	p->n = n;
	p->nr = 0;

	int m = n;

	while (0 == m % 4) { p->radix[p->nr++] = 4; m /= 4; }
	while (0 == m % 2) { p->radix[p->nr++] = 2; m /= 2; }

	for (int f = 3; m > 1; f += 2)
		while (0 == m % f) { p->radix[p->nr++] = f; m /= f; }

	p->tw = malloc(sizeof(cf) * (size_t)n);

	for (int k = 0; k < n; k++) {

		double a = -2. * M_PI * (double)k / (double)n;
		p->tw[k] = CMPLXF((float)cos(a), (float)sin(a));
	}
}

static void fft1d_plan_free(struct fft1d_plan* p)
{
	free(p->tw);
}

static inline cf twid(const struct fft1d_plan* p, long k, int sign)
{
	cf t = p->tw[k % p->n];
	return (sign < 0) ? t : conjf(t);
}

/* One length-n transform of x (work buffer y of the same size). */
static void fft1d(const struct fft1d_plan* p, int sign, cf* x, cf* y)
{
	int n = p->n;
	cf* src = x;
	cf* dst = y;
	long Ns = 1;

	for (int ir = 0; ir < p->nr; ir++) {

		int R = p->radix[ir];
		long nR = n / R;
		long tstep = n / (Ns * R);

		for (long j = 0; j < nR; j++) {

			long k = j % Ns;
			cf v[R];
			cf o[R];

			for (int r = 0; r < R; r++) {

				cf a = src[j + r * nR];
				v[r] = (0 == r || 0 == k) ? a : a * twid(p, (long)r * k * tstep, sign);
			}

			if (2 == R) {

				o[0] = v[0] + v[1];
				o[1] = v[0] - v[1];

			} else if (4 == R) {

				cf a0 = v[0] + v[2];
				cf a1 = v[0] - v[2];
				cf a2 = v[1] + v[3];
				cf a3 = v[1] - v[3];
				/* multiply a3 by -i (forward) or +i (backward) */
				cf a3r = (sign < 0) ? CMPLXF(cimagf(a3), -crealf(a3)) : CMPLXF(-cimagf(a3), crealf(a3));

				o[0] = a0 + a2;
				o[1] = a1 + a3r;
				o[2] = a0 - a2;
				o[3] = a1 - a3r;

			} else {

				for (int q = 0; q < R; q++) {

					cf s = 0.;

					for (int r = 0; r < R; r++)
						s += v[r] * twid(p, ((long)r * q % R) * nR, sign);

					o[q] = s;
				}
			}

			long idxD = (j / Ns) * Ns * R + k;

			for (int q = 0; q < R; q++)
				dst[idxD + q * Ns] = o[q];
		}

		cf* t = src;
		src = dst;
		dst = t;
		Ns *= R;
	}

	if (src != x)
		memcpy(x, src, sizeof(cf) * (size_t)n);
}

/* In-place FFT over dims[3] of each of `batch` contiguous volumes. */
void orc_fft(const long dims[3], long batch, int sign, cf* data)
{
	long N = dims[0] * dims[1] * dims[2];

	for (int a = 0; a < 3; a++) {

		if (1 >= dims[a])
			continue;

		int n = (int)dims[a];
		long stride = (0 == a) ? 1 : ((1 == a) ? dims[0] : dims[0] * dims[1]);
		long lines = N / n;

		struct fft1d_plan p;
		fft1d_plan_init(&p, n);

		#pragma omp parallel
		{
			cf* x = malloc(sizeof(cf) * (size_t)n);
			cf* y = malloc(sizeof(cf) * (size_t)n);

			#pragma omp for schedule(static)
			for (long l = 0; l < batch * lines; l++) {

				long b = l / lines;
				long li = l % lines;
				/* line base: decompose li over the other two dims */
				long base;

				if (0 == a)
					base = li * n;
				else if (1 == a)
					base = (li % dims[0]) + (li / dims[0]) * dims[0] * dims[1];
				else
					base = li;

				cf* d = data + b * N + base;

				for (int i = 0; i < n; i++)
					x[i] = d[i * stride];

				fft1d(&p, sign, x, y);

				for (int i = 0; i < n; i++)
					d[i * stride] = x[i];
			}

			free(x);
			free(y);
		}

		fft1d_plan_free(&p);
	}
}


/* ------------------------------------------------------------------------ */
/* fftmod, fftscale, linear phases                                          */
/* ------------------------------------------------------------------------ */

/* num/fft.c:64-69 */
static double fftmod_phase(long length, long j)
{
	long center1 = length / 2;
	double shift = (double)center1 / (double)length;
	return ((double)j - (double)center1 / 2.) * shift;
}

/* num/fft.c:71-90: exact values at quarter phases */
static double complex fftmod_value(double phase, bool inv)
{
	double rem = phase - floor(phase);
	double sgn = inv ? -1. : 1.;

	if (rem == 0.)
		return 1.;
	if (rem == 0.5)
		return -1.;
	if (rem == 0.25)
		return 1.i * sgn;
	if (rem == 0.75)
		return -1.i * sgn;

	return cexp(M_PI * 2.i * sgn * rem);
}

/* fftmod table over dims[3] (num/fft.c:92-157; all dims of size > 1 set) */
void orc_fftmod_table(const long dims[3], bool inv, cf* dst)
{
	#pragma omp parallel for collapse(2) schedule(static)
	for (long z = 0; z < dims[2]; z++)
		for (long y = 0; y < dims[1]; y++)
			for (long x = 0; x < dims[0]; x++) {

				long pos[3] = { x, y, z };
				double ph = 0.;

				for (int i = 0; i < 3; i++)
					if (dims[i] > 1)
						ph += fftmod_phase(dims[i], pos[i]);

				dst[x + dims[0] * (y + dims[1] * z)] = (cf)fftmod_value(ph, inv);
			}
}

/* Number of decomposition sub-grids and the shift of each:
 * noncart/nufft.c:123-145 (compute_factors, compute_shift). */
static int decomp_count(const long dims[3])
{
	int S = 1;

	for (int i = 0; i < 3; i++)
		if (dims[i] > 1)
			S *= 2;

	return S;
}

static void decomp_shift(const long dims[3], int idx, float shift[3])
{
	for (int i = 0; i < 3; i++) {

		int f = (dims[i] > 1) ? 2 : 1;
		shift[i] = -(float)(idx % f) / f;
		idx /= f;
	}
}

/* Linear phase table of sub-grid s, noncart/nufft.c:196-248 with
 * num/filter.c:241-265 (linear_phase, centered_gradient), then fftmod and
 * fftscale (nufft.c:945-946); roll folded in when roll != NULL
 * (nufft.c:933-942). */
void orc_linphase_table(const long dims[3], int s, const cf* roll, cf* dst)
{
	float sh[3];
	decomp_shift(dims, s, sh);

	double grad[3];
	double cn = 0.;

	for (int n = 0; n < 3; n++) {

		double shift = sh[n];

		if (dims[n] > 1)
			shift += (dims[n] / 2. - dims[n] / 2);

		grad[n] = 2. * M_PI * shift / (double)dims[n];
		cn -= grad[n] * (double)dims[n] / 2.;
	}

	long N = dims[0] * dims[1] * dims[2];
	double scale = 1. / sqrt((double)N);

	#pragma omp parallel for collapse(2) schedule(static)
	for (long z = 0; z < dims[2]; z++)
		for (long y = 0; y < dims[1]; y++)
			for (long x = 0; x < dims[0]; x++) {

				long pos[3] = { x, y, z };
				double ph = cn;
				double fm = 0.;

				for (int n = 0; n < 3; n++) {

					ph += grad[n] * (double)pos[n];

					if (dims[n] > 1)
						fm += fftmod_phase(dims[n], pos[n]);
				}

				long i = x + dims[0] * (y + dims[1] * z);
				double complex v = cexp(1.i * ph) * fftmod_value(fm, false) * scale;

				if (NULL != roll)
					v *= crealf(roll[i]);

				dst[i] = (cf)v;
			}
}


/* ------------------------------------------------------------------------ */
/* NUFFT operator                                                           */
/* ------------------------------------------------------------------------ */

struct orc_nufft {

	long dims[3];
	long N;		/* pixels per coil image */
	long M;		/* samples per coil */
	int C;		/* coils */
	int S;		/* sub-grids */
	float width;
	float os;
	double beta;
	bool periodic;
	bool toeplitz;

	cf* traj;	/* [M][3] */
	cf* weights;	/* [M] or NULL */
	cf* roll;	/* [N] (NULL when folded into linphase) */
	cf* fftmod;	/* [N] */
	cf* linphase;	/* [S][N] */
	cf* psf;	/* [S][N] (Toeplitz) */
	cf* grid;	/* work: [S][C][N] */

	float table[KB_SIZE + 1];
};

typedef struct orc_nufft orc_nufft_t;

static struct orc_grid_conf decomp_conf(const orc_nufft_t* nf, int s)
{
	/* noncart/nufft.c:147-188: width / 2, os = 1, shift per sub-grid,
	 * odd-size centre correction */
	struct orc_grid_conf g = { .os = 1., .width = nf->width / 2., .periodic = nf->periodic };

	decomp_shift(nf->dims, s, g.shift);

	for (int i = 0; i < 3; i++)
		if (nf->dims[i] > 1)
			g.shift[i] += (nf->dims[i] / 2.0 - nf->dims[i] / 2) / g.os;

	return g;
}

void orc_nufft_compute_psf(const long dims[3], long M, const cf* traj, const cf* weights, cf* psf);

/* noncart/nufft.c:779-999 (tables) and :1002-1159 (trajectory, PSF). */
orc_nufft_t* orc_nufft_create(const long dims[3], long M, const cf* traj, const cf* weights,
			int C, bool toeplitz, bool periodic, float width, float os)
{
	if (2.f != os)
		return NULL;

	orc_nufft_t* nf = calloc(1, sizeof(*nf));

	for (int i = 0; i < 3; i++)
		nf->dims[i] = dims[i];

	nf->N = dims[0] * dims[1] * dims[2];
	nf->M = M;
	nf->C = C;
	nf->S = decomp_count(dims);
	nf->width = width;
	nf->os = os;
	/* BART builds the KB table on the first kb_init call, which comes from
	 * rolloff_correction with beta passed as float (nufft.c:898,
	 * grid.c:537-540); later calls with the double beta pass its 1e-6
	 * check (grid.c:58-59).  So table, I0(beta) and rolloff all use the
	 * float-rounded beta. */
	nf->beta = (double)(float)orc_calc_beta(os, width);
	nf->periodic = periodic;
	nf->toeplitz = toeplitz;

	orc_kb_table(nf->beta, nf->table);

	nf->traj = malloc(sizeof(cf) * 3 * (size_t)M);
	memcpy(nf->traj, traj, sizeof(cf) * 3 * (size_t)M);

	if (NULL != weights) {

		nf->weights = malloc(sizeof(cf) * (size_t)M);
		memcpy(nf->weights, weights, sizeof(cf) * (size_t)M);
	}

	nf->fftmod = malloc(sizeof(cf) * (size_t)nf->N);
	orc_fftmod_table(dims, false, nf->fftmod);

	nf->roll = malloc(sizeof(cf) * (size_t)nf->N);
	orc_rolloff(os, width, nf->beta, dims, nf->roll);

	nf->linphase = malloc(sizeof(cf) * (size_t)(nf->S * nf->N));

	for (int s = 0; s < nf->S; s++)
		orc_linphase_table(dims, s, toeplitz ? NULL : nf->roll, nf->linphase + s * nf->N);

	if (!toeplitz) {

		free(nf->roll);
		nf->roll = NULL;
	}

	nf->grid = malloc(sizeof(cf) * (size_t)(nf->S * C) * (size_t)nf->N);

	if (toeplitz) {

		nf->psf = malloc(sizeof(cf) * (size_t)(nf->S * nf->N));
		orc_nufft_compute_psf(dims, M, traj, weights, nf->psf);
	}

	return nf;
}

void orc_nufft_free(orc_nufft_t* nf)
{
	if (NULL == nf)
		return;

	free(nf->traj);
	free(nf->weights);
	free(nf->roll);
	free(nf->fftmod);
	free(nf->linphase);
	free(nf->psf);
	free(nf->grid);
	free(nf);
}

/* Forward: coil images -> k-space, noncart/nufft.c:1388-1458 (decomp). */
void orc_nufft_forward(const orc_nufft_t* nf, cf* ksp, const cf* cim)
{
	long N = nf->N;
	long M = nf->M;
	int C = nf->C;
	int S = nf->S;
	cf* grid = nf->grid;

	#pragma omp parallel for collapse(2) schedule(static)
	for (int s = 0; s < S; s++)
		for (int c = 0; c < C; c++)
			for (long i = 0; i < N; i++) {

				cf v = cim[i + c * N];

				if (NULL != nf->roll)
					v = v * nf->roll[i];

				grid[i + N * (c + C * (long)s)] = v * nf->linphase[i + s * N];
			}

	orc_fft(nf->dims, (long)S * C, -1, grid);

	#pragma omp parallel for schedule(static)
	for (long j = 0; j < (long)S * C; j++)
		for (long i = 0; i < N; i++)
			grid[i + j * N] *= nf->fftmod[i];

	memset(ksp, 0, sizeof(cf) * (size_t)(M * C));

	for (int s = 0; s < S; s++) {

		struct orc_grid_conf g = decomp_conf(nf, s);
		orc_gridH(&g, nf->table, M, nf->traj, C, ksp, M, nf->dims, grid + (long)s * C * N, N);
	}

	if (NULL != nf->weights)
		for (int c = 0; c < C; c++)
			for (long m = 0; m < M; m++)
				ksp[m + c * M] *= nf->weights[m];
}

/* Adjoint: k-space -> coil images, noncart/nufft.c:1462-1540 (decomp). */
void orc_nufft_adjoint(const orc_nufft_t* nf, cf* cim, const cf* ksp)
{
	long N = nf->N;
	long M = nf->M;
	int C = nf->C;
	int S = nf->S;
	cf* grid = nf->grid;
	cf* wk = NULL;

	if (NULL != nf->weights) {

		wk = malloc(sizeof(cf) * (size_t)(M * C));

		for (int c = 0; c < C; c++)
			for (long m = 0; m < M; m++)
				wk[m + c * M] = ksp[m + c * M] * conjf(nf->weights[m]);

		ksp = wk;
	}

	memset(grid, 0, sizeof(cf) * (size_t)(S * C) * (size_t)N);

	for (int s = 0; s < S; s++) {

		struct orc_grid_conf g = decomp_conf(nf, s);
		orc_grid(&g, nf->table, M, nf->traj, nf->dims, C, grid + (long)s * C * N, N, ksp, M);
	}

	free(wk);

	#pragma omp parallel for schedule(static)
	for (long j = 0; j < (long)S * C; j++)
		for (long i = 0; i < N; i++)
			grid[i + j * N] *= conjf(nf->fftmod[i]);

	orc_fft(nf->dims, (long)S * C, +1, grid);

	#pragma omp parallel for schedule(static)
	for (long ci = 0; ci < (long)C * N; ci++) {

		long c = ci / N;
		long i = ci % N;
		cf acc = 0.;

		for (int s = 0; s < S; s++)
			acc += grid[i + N * (c + C * (long)s)] * conjf(nf->linphase[i + s * N]);

		if (NULL != nf->roll)
			acc *= nf->roll[i];

		cim[ci] = acc;
	}
}

/* Toeplitz normal operator, noncart/nufft.c:1546-1625: for every coil
 * sum_s conj(L_s) IFFT(P_s FFT(L_s x)). */
void orc_nufft_normal(const orc_nufft_t* nf, cf* dst, const cf* src)
{
	long N = nf->N;
	int C = nf->C;
	int S = nf->S;
	cf* grid = nf->grid;

	#pragma omp parallel for collapse(2) schedule(static)
	for (int s = 0; s < S; s++)
		for (int c = 0; c < C; c++)
			for (long i = 0; i < N; i++)
				grid[i + N * (c + C * (long)s)] = src[i + c * N] * nf->linphase[i + s * N];

	orc_fft(nf->dims, (long)S * C, -1, grid);

	#pragma omp parallel for collapse(2) schedule(static)
	for (int s = 0; s < S; s++)
		for (int c = 0; c < C; c++)
			for (long i = 0; i < N; i++)
				grid[i + N * (c + C * (long)s)] *= nf->psf[i + s * N];

	orc_fft(nf->dims, (long)S * C, +1, grid);

	#pragma omp parallel for schedule(static)
	for (long ci = 0; ci < (long)C * N; ci++) {

		long c = ci / N;
		long i = ci % N;
		cf acc = 0.;

		for (int s = 0; s < S; s++)
			acc += grid[i + N * (c + C * (long)s)] * conjf(nf->linphase[i + s * N]);

		dst[ci] = acc;
	}
}

/* Toeplitz normal operator with a subspace basis, noncart/nufft.c:1546-1625
 * with the K x K PSF of nufft_set_traj (:1066-1079): nf has C K channels
 * (coil c, coefficient k at c + C k); psfB is [K K][S][N], entry k K + k'
 * multiplying coefficient k' into coefficient k (md_ztenmul over the
 * coefficient dims, the sum over k' in order). */
void orc_nufft_normal_basis(const orc_nufft_t* nf, int K, const cf* psfB, cf* dst, const cf* src)
{
	long N = nf->N;
	int CK = nf->C;
	int C = CK / K;
	int S = nf->S;
	cf* grid = nf->grid;

	#pragma omp parallel for collapse(2) schedule(static)
	for (int s = 0; s < S; s++)
		for (int c = 0; c < CK; c++)
			for (long i = 0; i < N; i++)
				grid[i + N * (c + CK * (long)s)] = src[i + c * N] * nf->linphase[i + s * N];

	orc_fft(nf->dims, (long)S * CK, -1, grid);

	#pragma omp parallel for collapse(2) schedule(static)
	for (int s = 0; s < S; s++)
		for (int c = 0; c < C; c++) {

			cf v[64];

			for (long i = 0; i < N; i++) {

				for (int k = 0; k < K; k++)
					v[k] = grid[i + N * (c + C * k + CK * (long)s)];

				for (int k = 0; k < K; k++) {

					cf acc = 0.;

					for (int q = 0; q < K; q++)
						acc += psfB[i + N * (s + S * ((long)k * K + q))] * v[q];

					grid[i + N * (c + C * k + CK * (long)s)] = acc;
				}
			}
		}

	orc_fft(nf->dims, (long)S * CK, +1, grid);

	#pragma omp parallel for schedule(static)
	for (long ci = 0; ci < (long)CK * N; ci++) {

		long c = ci / N;
		long i = ci % N;
		cf acc = 0.;

		for (int s = 0; s < S; s++)
			acc += grid[i + N * (c + CK * (long)s)] * conjf(nf->linphase[i + s * N]);

		dst[ci] = acc;
	}
}

/* PSF, noncart/nufft.c:727-776 (compute_psf2) with :535-586
 * (compute_psf_int): adjoint NUFFT of |w|^2 (ones without weights) on the
 * 2x image with 2*traj and periodic gridding, centred unitary FFT over
 * the 2x grid (num/fft.c:471-475), then even/odd decomposition
 * (num/shuffle.c:84-130).  psf is [S][N]. */
void orc_nufft_compute_psf_vals(const long dims[3], long M, const cf* traj, const cf* weights, const cf* vals,
				cf* psf);

void orc_nufft_compute_psf(const long dims[3], long M, const cf* traj, const cf* weights, cf* psf)
{
	orc_nufft_compute_psf_vals(dims, M, traj, weights, NULL, psf);
}

/* The same with the k-space vals[M] in place of the ones: one channel of the
 * PSF of a subspace basis, whose adjoint NUFFT takes |w|^2 conj(sqr_basis)
 * (compute_square_basis, noncart/nufft.c:470-505; nufft_apply_adjoint
 * :1483-1487), vals = conj(B[t][k]) B[t][k'] for the PSF entry (k, k'). */
void orc_nufft_compute_psf_vals(const long dims[3], long M, const cf* traj, const cf* weights, const cf* vals,
				cf* psf)
{
	long dims2[3];

	for (int i = 0; i < 3; i++)
		dims2[i] = (1 == dims[i]) ? 1 : 2 * dims[i];

	long N2 = dims2[0] * dims2[1] * dims2[2];

	cf* traj2 = malloc(sizeof(cf) * 3 * (size_t)M);

	for (long i = 0; i < 3 * M; i++)
		traj2[i] = traj[i] * 2.f;

	cf* sqw = NULL;

	if (NULL != weights) {

		sqw = malloc(sizeof(cf) * (size_t)M);

		for (long m = 0; m < M; m++)
			sqw[m] = weights[m] * conjf(weights[m]);
	}

	orc_nufft_t* nf2 = orc_nufft_create(dims2, M, traj2, sqw, 1, false, true, 6., 2.);

	/* compute_psf_nufft_conf (nufft.c:519-532) disables precomputation:
	 * roll is applied separately, not folded into the linear phases */
	nf2->roll = malloc(sizeof(cf) * (size_t)N2);
	orc_rolloff(nf2->os, nf2->width, nf2->beta, dims2, nf2->roll);

	for (int s = 0; s < nf2->S; s++)
		orc_linphase_table(dims2, s, NULL, nf2->linphase + s * N2);

	cf* ones = malloc(sizeof(cf) * (size_t)M);

	for (long m = 0; m < M; m++)
		ones[m] = (NULL != vals) ? vals[m] : 1.;

	cf* psft = malloc(sizeof(cf) * (size_t)N2);

	orc_nufft_adjoint(nf2, psft, ones);

	/* fftuc = fftmod, fft, fftmod, fftscale */
	cf* fm = malloc(sizeof(cf) * (size_t)N2);
	orc_fftmod_table(dims2, false, fm);

	for (long i = 0; i < N2; i++)
		psft[i] *= fm[i];

	orc_fft(dims2, 1, -1, psft);

	float scale = 1.f / sqrtf((float)N2);

	for (long i = 0; i < N2; i++)
		psft[i] = psft[i] * fm[i] * scale;

	/* md_decompose: psf[s][j] = psft[2 j + r(s)] */
	int S = decomp_count(dims);
	long N = dims[0] * dims[1] * dims[2];

	for (int s = 0; s < S; s++) {

		int r[3];
		int idx = s;

		for (int i = 0; i < 3; i++) {

			int f = (dims[i] > 1) ? 2 : 1;
			r[i] = idx % f;
			idx /= f;
		}

		for (long z = 0; z < dims[2]; z++)
			for (long y = 0; y < dims[1]; y++)
				for (long x = 0; x < dims[0]; x++) {

					long xx = (dims[0] > 1) ? 2 * x + r[0] : x;
					long yy = (dims[1] > 1) ? 2 * y + r[1] : y;
					long zz = (dims[2] > 1) ? 2 * z + r[2] : z;

					psf[s * N + x + dims[0] * (y + dims[1] * z)] = psft[xx + dims2[0] * (yy + dims2[1] * zz)];
				}
	}

	free(fm);
	free(psft);
	free(ones);
	free(sqw);
	free(traj2);
	orc_nufft_free(nf2);
}

/* The gridding step of the PSF above (the adjoint NUFFT of |w|^2 on the 2x
 * image with 2*traj, periodic, sub-grid s of its decomposition) with the
 * kernel products and the sums in double: the FP64 reference for the
 * rounding of the PSF's accumulation (not a BART function; the remaining
 * steps of compute_psf2 are applied in complex128 by the caller).  grid is
 * [dims2] complex128, zeroed here. */
void orc_psf_grid_fp64(const long dims[3], long M, const cf* traj, const cf* weights, int s,
		double complex* grid)
{
	long dims2[3];

	for (int i = 0; i < 3; i++)
		dims2[i] = (1 == dims[i]) ? 1 : 2 * dims[i];

	long N2 = dims2[0] * dims2[1] * dims2[2];

	cf* traj2 = malloc(sizeof(cf) * 3 * (size_t)M);

	for (long i = 0; i < 3 * M; i++)
		traj2[i] = traj[i] * 2.f;

	cf* val = malloc(sizeof(cf) * (size_t)M);

	for (long m = 0; m < M; m++)
		val[m] = (NULL != weights) ? weights[m] * conjf(weights[m]) : 1.f;

	float table[KB_SIZE + 1];
	orc_kb_table((double)(float)orc_calc_beta(2.f, 6.f), table);

	/* decomp_conf of the 2x operator: width 6 / 2, os 1, periodic */
	struct orc_grid_conf g = { .os = 1., .width = 3., .periodic = true };

	decomp_shift(dims2, s, g.shift);

	for (int i = 0; i < 3; i++)
		if (dims2[i] > 1)
			g.shift[i] += (dims2[i] / 2.0 - dims2[i] / 2) / g.os;

	memset(grid, 0, sizeof(double complex) * (size_t)N2);
	grid_core(&g, table, M, traj2, dims2, 1, NULL, grid, N2, val, M);

	free(val);
	free(traj2);
}

/* The gridding step of the adjoint above for sub-grid s (k-space times
 * conj(weights), noncart/nufft.c:1462-1540) with the kernel products and
 * the sums in double: the FP64 reference for the rounding of the adjoint's
 * accumulation (not a BART function; the caller applies conj(fftmod), the
 * FFT and the linear phases in complex128).  grid is [C][N] complex128,
 * zeroed here. */
void orc_nufft_grid_fp64(const orc_nufft_t* nf, int s, const cf* ksp, double complex* grid)
{
	long M = nf->M;
	int C = nf->C;
	cf* wk = NULL;

	if (NULL != nf->weights) {

		wk = malloc(sizeof(cf) * (size_t)(M * C));

		for (int c = 0; c < C; c++)
			for (long m = 0; m < M; m++)
				wk[m + c * M] = ksp[m + c * M] * conjf(nf->weights[m]);

		ksp = wk;
	}

	memset(grid, 0, sizeof(double complex) * (size_t)C * (size_t)nf->N);

	struct orc_grid_conf g = decomp_conf(nf, s);
	grid_core(&g, nf->table, M, nf->traj, nf->dims, C, NULL, grid, nf->N, ksp, M);

	free(wk);
}

void orc_nufft_get_psf(const orc_nufft_t* nf, cf* psf)
{
	memcpy(psf, nf->psf, sizeof(cf) * (size_t)(nf->S * nf->N));
}

void orc_nufft_set_psf(orc_nufft_t* nf, const cf* psf)
{
	memcpy(nf->psf, psf, sizeof(cf) * (size_t)(nf->S * nf->N));
}

void orc_nufft_get_linphase(const orc_nufft_t* nf, cf* lph)
{
	memcpy(lph, nf->linphase, sizeof(cf) * (size_t)(nf->S * nf->N));
}

int orc_nufft_subgrids(const orc_nufft_t* nf)
{
	return nf->S;
}

double orc_nufft_beta(const orc_nufft_t* nf)
{
	return nf->beta;
}


/* ------------------------------------------------------------------------ */
/* SENSE model and CG                                                       */
/* ------------------------------------------------------------------------ */

/* S: coil_img[c] = maps[c] * img (linops/fmac.c:63-86, sense/model.c:93) */
static void sense_fwd(long N, int C, const cf* maps, cf* cim, const cf* img)
{
	#pragma omp parallel for schedule(static)
	for (long ci = 0; ci < C * N; ci++)
		cim[ci] = maps[ci] * img[ci % N];
}

/* S^H: img = sum_c conj(maps[c]) coil_img[c] (md_ztenmulc) */
static void sense_adj(long N, int C, const cf* maps, cf* img, const cf* cim)
{
	#pragma omp parallel for schedule(static)
	for (long i = 0; i < N; i++) {

		cf acc = 0.;

		for (int c = 0; c < C; c++)
			acc += cim[i + c * N] * conjf(maps[i + c * N]);

		img[i] = acc;
	}
}

/* A = NUFFT S (sense/modelnc.c:28-67) */
void orc_sense_forward(const orc_nufft_t* nf, const cf* maps, cf* ksp, const cf* img)
{
	cf* cim = malloc(sizeof(cf) * (size_t)(nf->C * nf->N));
	sense_fwd(nf->N, nf->C, maps, cim, img);
	orc_nufft_forward(nf, ksp, cim);
	free(cim);
}

void orc_sense_adjoint(const orc_nufft_t* nf, const cf* maps, cf* img, const cf* ksp)
{
	cf* cim = malloc(sizeof(cf) * (size_t)(nf->C * nf->N));
	orc_nufft_adjoint(nf, cim, ksp);
	sense_adj(nf->N, nf->C, maps, img, cim);
	free(cim);
}

/* A^H A via Toeplitz (linops/linop.c:606-637 chain normal) or, without
 * a PSF, via adjoint(forward(.)) */
void orc_sense_normal(const orc_nufft_t* nf, const cf* maps, cf* dst, const cf* src)
{
	cf* cim = malloc(sizeof(cf) * (size_t)(nf->C * nf->N));
	cf* cim2 = malloc(sizeof(cf) * (size_t)(nf->C * nf->N));

	sense_fwd(nf->N, nf->C, maps, cim, src);

	if (nf->toeplitz) {

		orc_nufft_normal(nf, cim2, cim);

	} else {

		cf* ksp = malloc(sizeof(cf) * (size_t)(nf->C * nf->M));
		orc_nufft_forward(nf, ksp, cim);
		orc_nufft_adjoint(nf, cim2, ksp);
		free(ksp);
	}

	sense_adj(nf->N, nf->C, maps, dst, cim2);

	free(cim);
	free(cim2);
}

/* num/vecops.c:89-124: float products accumulated in double */
static double vdot(long n, const float* a, const float* b)
{
	double res = 0.;

	for (long i = 0; i < n; i++)
		res += a[i] * b[i];

	return res;
}

static double vnorm(long n, const float* a)
{
	return sqrt(vdot(n, a, a));
}

static void vaxpy(long n, float* y, float alpha, const float* x)
{
	for (long i = 0; i < n; i++)
		y[i] += alpha * x[i];
}

static void vxpay(long n, float beta, float* y, const float* x)
{
	for (long i = 0; i < n; i++)
		y[i] = beta * y[i] + x[i];
}

/* CG on (A^H A + lambda I) x = b, iter/italgos.c:497-580 with the stopping
 * rule of iter/iter2.c:227-270 (eps = tol * ||b||).  x holds the initial
 * guess on entry.  Returns the number of iterations; rsnew_trace (may be
 * NULL) receives sqrt(rsnew) after each iteration. */
int orc_cgsense(const orc_nufft_t* nf, const cf* maps, float lambda, int maxiter, float tol,
		cf* x, const cf* b, float* resid_trace, float* final_resid)
{
	long N2 = 2 * nf->N;
	float* xf = (float*)x;
	const float* bf = (const float*)b;

	float* r = malloc(sizeof(float) * (size_t)N2);
	float* p = malloc(sizeof(float) * (size_t)N2);
	float* Ap = malloc(sizeof(float) * (size_t)N2);

	float eps = (float)vnorm(N2, bf) * tol;

	int i = 0;
	float rsnew = 0.;

	if (0. == vnorm(N2, bf))	/* checkeps, iter/iter2.c:65-80 */
		goto out;

	orc_sense_normal(nf, maps, (cf*)r, x);
	vaxpy(N2, r, lambda, xf);
	vxpay(N2, -1., r, bf);
	memcpy(p, r, sizeof(float) * (size_t)N2);

	float rsold = (float)pow(vnorm(N2, r), 2.);
	rsnew = rsold;

	float eps_squared = (float)pow(eps, 2.);

	if (0. == rsold)
		goto out;

	for (i = 0; i < maxiter; i++) {

		orc_sense_normal(nf, maps, (cf*)Ap, (const cf*)p);
		vaxpy(N2, Ap, lambda, p);

		float pAp = (float)vdot(N2, p, Ap);

		if (0. == pAp)
			break;

		float alpha = rsold / pAp;

		vaxpy(N2, xf, +alpha, p);
		vaxpy(N2, r, -alpha, Ap);

		rsnew = (float)pow(vnorm(N2, r), 2.);

		if (NULL != resid_trace)
			resid_trace[i] = sqrtf(rsnew);

		float beta = rsnew / rsold;

		rsold = rsnew;

		if (rsnew <= eps_squared) {

			i++;
			break;
		}

		vxpay(N2, beta, p, r);
	}

out:
	if (NULL != final_resid)
		*final_resid = sqrtf(rsnew);

	free(r);
	free(p);
	free(Ap);

	return i;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
	omp_set_num_threads(n);
#else
	(void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* The chained NUFFT for os != 2, noncart/nufft_chain.c:192-244 (selected   */
/* by nufft_create2, noncart/nufft.c:1184-1199): rolloff into the centre of */
/* the os-times grid (linop_kb_rolloff_create, :60-118, rolloff of          */
/* grid.c:520-556), centred unitary FFT (linop_fftc_create,                 */
/* linops/someops.c:1825-1879: fftmod(N^-1/4) x FFT x fftmod(N^-1/4)), KB   */
/* interpolation on that grid (grid2H with os, width, beta =               */
/* calc_beta(os, width)), and the weights times os^(d/2).  No Toeplitz.     */
/* ------------------------------------------------------------------------ */

struct orc_chain {

	long dims[3], odims[3], off[3];
	long N, N2, M;
	int C;
	float scale;		/* os^(d/2) */
	struct orc_grid_conf conf;
	cf* traj;
	cf* tmp;		/* [M] weights x scale, or NULL */
	cf* roll;		/* [N] */
	cf* fma;		/* [N2] fftmod x N2^-1/4 */
	cf* grid;		/* [C][N2] */
	float table[KB_SIZE + 1];
};

typedef struct orc_chain orc_chain_t;

/* pad = false: the non-decomposed NUFFT of nufft_create2 with decomp =
 * false (noncart/nufft.c:898-935, 1432-1441): the image is the
 * oversampled grid itself (no padding), the rolloff is taken with os = 1
 * over it, the positions with os, and the scale 2^(d/2) folded into the
 * linear phase is the os^(d/2) of the chain. */
orc_chain_t* orc_chain_create(const long dims[3], long M, const cf* traj, const cf* weights, int C, bool periodic,
			      float width, float os, bool pad)
{
	orc_chain_t* ch = calloc(1, sizeof(orc_chain_t));
	int d = 0;

	ch->N = ch->N2 = 1;

	for (int i = 0; i < 3; i++) {

		ch->dims[i] = dims[i];
		ch->odims[i] = (dims[i] > 1 && pad) ? lround(os * dims[i]) : dims[i];
		ch->off[i] = labs(ch->odims[i] / 2 - dims[i] / 2);
		ch->N *= dims[i];
		ch->N2 *= ch->odims[i];
		d += (dims[i] > 1);
	}

	ch->M = M;
	ch->C = C;
	ch->conf = (struct orc_grid_conf){ .os = os, .width = width, .periodic = periodic };

	/* the table and the rolloff use beta rounded to float, as the first
	 * kb_init call of the forward (the rolloff's) makes it */
	double beta = (double)(float)orc_calc_beta(os, width);
	orc_kb_table(beta, ch->table);

	ch->roll = malloc(sizeof(cf) * (size_t)ch->N);
	orc_rolloff(pad ? os : 1.f, width, beta, dims, ch->roll);

	ch->fma = malloc(sizeof(cf) * (size_t)ch->N2);
	orc_fftmod_table(ch->odims, false, ch->fma);
	float s4 = 1.f / sqrtf(sqrtf((float)ch->N2));

	for (long i = 0; i < ch->N2; i++)
		ch->fma[i] *= s4;

	ch->scale = powf(os, d / 2.f);

	ch->traj = malloc(sizeof(cf) * 3 * (size_t)M);
	memcpy(ch->traj, traj, sizeof(cf) * 3 * (size_t)M);

	if (NULL != weights) {

		ch->tmp = malloc(sizeof(cf) * (size_t)M);

		for (long m = 0; m < M; m++)
			ch->tmp[m] = weights[m] * ch->scale;
	}

	ch->grid = malloc(sizeof(cf) * (size_t)C * (size_t)ch->N2);
	return ch;
}

void orc_chain_free(orc_chain_t* ch)
{
	if (NULL == ch)
		return;

	free(ch->traj);
	free(ch->tmp);
	free(ch->roll);
	free(ch->fma);
	free(ch->grid);
	free(ch);
}

void orc_chain_forward(orc_chain_t* ch, cf* ksp, const cf* cim)
{
	long N = ch->N, N2 = ch->N2;
	int C = ch->C;
	cf* g = ch->grid;

	memset(g, 0, sizeof(cf) * (size_t)C * (size_t)N2);

	for (int c = 0; c < C; c++)
		for (long z = 0; z < ch->dims[2]; z++)
			for (long y = 0; y < ch->dims[1]; y++)
				for (long x = 0; x < ch->dims[0]; x++) {

					long i = x + ch->dims[0] * (y + ch->dims[1] * z);
					long o = (x + ch->off[0]) + ch->odims[0] * ((y + ch->off[1]) + ch->odims[1] * (z + ch->off[2]));
					g[c * N2 + o] = crealf(ch->roll[i]) * cim[c * N + i];
				}

	for (int c = 0; c < C; c++)
		for (long o = 0; o < N2; o++)
			g[c * N2 + o] *= ch->fma[o];

	orc_fft(ch->odims, C, -1, g);

	for (int c = 0; c < C; c++)
		for (long o = 0; o < N2; o++)
			g[c * N2 + o] *= ch->fma[o];

	memset(ksp, 0, sizeof(cf) * (size_t)C * (size_t)ch->M);
	orc_gridH(&ch->conf, ch->table, ch->M, ch->traj, C, ksp, ch->M, ch->odims, g, N2);

	for (int c = 0; c < C; c++)
		for (long m = 0; m < ch->M; m++)
			ksp[c * ch->M + m] *= (NULL != ch->tmp) ? ch->tmp[m] : ch->scale;
}

void orc_chain_adjoint(orc_chain_t* ch, cf* cim, const cf* ksp)
{
	long N = ch->N, N2 = ch->N2, M = ch->M;
	int C = ch->C;
	cf* g = ch->grid;
	cf* k2 = malloc(sizeof(cf) * (size_t)C * (size_t)M);

	for (int c = 0; c < C; c++)
		for (long m = 0; m < M; m++)
			k2[c * M + m] = ksp[c * M + m] * ((NULL != ch->tmp) ? conjf(ch->tmp[m]) : ch->scale);

	memset(g, 0, sizeof(cf) * (size_t)C * (size_t)N2);
	orc_grid(&ch->conf, ch->table, M, ch->traj, ch->odims, C, g, N2, k2, M);
	free(k2);

	for (int c = 0; c < C; c++)
		for (long o = 0; o < N2; o++)
			g[c * N2 + o] *= conjf(ch->fma[o]);

	orc_fft(ch->odims, C, +1, g);

	for (int c = 0; c < C; c++)
		for (long o = 0; o < N2; o++)
			g[c * N2 + o] *= conjf(ch->fma[o]);

	for (int c = 0; c < C; c++)
		for (long z = 0; z < ch->dims[2]; z++)
			for (long y = 0; y < ch->dims[1]; y++)
				for (long x = 0; x < ch->dims[0]; x++) {

					long i = x + ch->dims[0] * (y + ch->dims[1] * z);
					long o = (x + ch->off[0]) + ch->odims[0] * ((y + ch->off[1]) + ch->odims[1] * (z + ch->off[2]));
					cim[c * N + i] = crealf(ch->roll[i]) * g[c * N2 + o];
				}
}
