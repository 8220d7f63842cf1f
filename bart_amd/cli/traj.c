/*
 * traj.c -- "bart traj" for the trajectories of this path (src/traj.c:52-460,
 * noncart/traj.c:61-270): Cartesian, radial, golden-angle radial (-G, tiny
 * golden angles -s, --double-base), full circle (-D), rotation (-R),
 * oversampling (-o) and the 3-D kooshball (-3).  Multi-echo, SMS, turns,
 * gradient delays, custom angles and rational approximations are rejected.
 *
 *   traj [-x X] [-y Y] [-r] [-G] [-s n] [-D] [--double-base] [-H] [-o over]
 *        [-R deg] [-3] [-c] output
 */
#include "tools.h"
#include "cfl.h"

#include <complex.h>
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static void usage(void)
{
	fprintf(stderr,
		"Usage: traj [-x X] [-y Y] [-r] [-G] [-s n] [-D] [--double-base] [-H] [-o over] [-R phi] [-3] [-c] <output>\n"
		"\n"
		"Computes k-space trajectories.\n");
}

// noncart/traj.c:154-181
static double golden_angle(int tiny)
{
	if (1 == tiny)
		return M_PI * (2. - (3. - sqrt(5.))) / 2.;

	double ratio = (sqrt(5.) + 1.) / 2.;
	return M_PI / (ratio + tiny - 1.);
}

// noncart/traj.c:40-66: read direction of angles (phi, psi), x and y swapped
static void read_dir(float dir[3], float phi, float psi)
{
	dir[1] = cosf(phi) * cosf(psi);
	dir[0] = sinf(phi) * cosf(psi);
	dir[2] = sinf(psi);
}

int main_traj(int argc, char* argv[])
{
	int X = -1, Y = -1, Z = 1;
	float over = 1.f;
	float rot = 0.f;
	bool radial = false, golden = false, full_circle = false, double_base = false;
	bool half_circle_gold = false, d3d = false, asym = false;
	int tiny = 0;

	enum { O_DOUBLEBASE = 256, O_UNSUP };

	static const struct option lopts[] = {
		{ "double-base", no_argument, NULL, O_DOUBLEBASE },
		{ "mems-legacy", no_argument, NULL, O_UNSUP },
		{ "z-us", required_argument, NULL, O_UNSUP },
		{ "raga-inc", required_argument, NULL, O_UNSUP },
		{ NULL, 0, NULL, 0 },
	};

	optind = 1;
	int c, li = -1;

	while (-1 != (c = getopt_long(argc, argv, "x:y:z:d:e:a:t:m:lgrGHs:ADo:R:q:Q:O3cEC:V:h", lopts, &li))) {

		switch (c) {
		case 'x': X = parse_int(optarg, "-x"); break;
		case 'y': Y = parse_int(optarg, "-y"); break;
		case 'z': Z = parse_int(optarg, "-z"); break;
		case 'r': radial = true; break;
		case 'G': golden = true; break;
		case 'H': half_circle_gold = true; break;
		case 's': tiny = parse_int(optarg, "-s"); break;
		case 'D': full_circle = true; break;
		case 'o': over = parse_float(optarg, "-o"); break;
		case 'R': rot = parse_float(optarg, "-R"); break;
		case '3': d3d = true; break;
		case 'c': asym = true; break;
		case O_DOUBLEBASE: double_base = true; break;
		case O_UNSUP: unsupported_option("traj", lopts[li].name); break;
		case 'h': usage(); return 0;
		case '?': usage(); return 1;
		default: {
			char o[3] = { '-', (char)c, 0 };
			unsupported_option("traj", o);
		}
		}

		li = -1;
	}

	if (1 != argc - optind) {
		usage();
		cli_error("traj: expected 1 file argument");
	}

	if (-1 == Y)
		Y = 128;

	if (-1 == X)
		X = 128;

	if (over <= 0.f)
		cli_error("Oversampling factor must be positive.\n");

	X = (int)((float)X * over);
	const int D = X;

	if (half_circle_gold) {
		golden = true;
		if (full_circle)
			cli_error("Invalid options. Full-circle or half-circle sampling?\n");
	}

	if (tiny >= 1)
		golden = true;

	if (golden) {
		radial = true;
		if (0 == tiny)
			tiny = 1;
	} else if (full_circle) {
		radial = true;
	}

	if (radial && Z > 1 && !d3d)
		cli_error("traj: -z with radial trajectories is not supported");

	long dims[DIMS];

	for (int i = 0; i < DIMS; i++)
		dims[i] = 1;

	dims[0] = 3;
	dims[1] = X;
	dims[2] = (long)Y * Z;

	complex float* samples = calloc((size_t)cfl_size(DIMS, dims), sizeof(complex float));

	// calc_base_angles (noncart/traj.c:191-270): angle between spokes
	double ga = golden_angle(tiny);

	if (double_base)
		ga *= 2.;

	const double angle_s = golden ? ga : (M_PI / Y) * (full_circle ? 2 : 1);

	long p = 0;
	double phin1 = 0.;

	for (long s = 0; s < dims[2]; s++) {

		const int j = (int)(s % Y);
		const int k = (int)(s / Y);

		for (int i = 0; i < X; i++, p++) {

			if (radial) {

				const int sample = i + D - X;
				const double read = (float)sample + (asym ? 0 : 0.5) - (float)D / 2.;

				double angle = (double)j * angle_s;

				if (half_circle_gold)
					angle = fmod(angle, M_PI);

				angle += M_PI * rot / 180.;

				double angle2 = 0.;

				if (d3d) {

					// Saff EB, Kuijlaars ABJ. Distributing many points on a
					// sphere. Math Intelligencer 1997;19:11 (src/traj.c:388-414)
					const int Y2 = full_circle ? Y : 2 * Y;
					const double hn = -1.0 + (double)(2 * j) / (Y2 - 1);

					if ((j + 1 == Y) || (0 == j))
						angle = 0;
					else
						angle = fmod(phin1 + 3.6 / sqrt(Y2 * (1.0 - hn * hn)), 2. * M_PI);

					if (i + 1 == X)
						phin1 = angle;

					angle2 = acos(hn) - M_PI / 2.;
				}

				float dir[3];
				read_dir(dir, (float)angle, (float)angle2);

				samples[p * 3 + 0] = (0.f + read * dir[0]) / over;
				samples[p * 3 + 1] = (0.f + read * dir[1]) / over;
				samples[p * 3 + 2] = (0.f + read * dir[2]) / over;

			} else {

				const double x = (i - X / 2) / over;
				const double y = (j - Y / 2);
				const double z = (k - Z / 2);
				const double a = -rot / 180. * M_PI;

				samples[p * 3 + 0] = x * cos(a) + -y * sin(a);
				samples[p * 3 + 1] = x * sin(a) + y * cos(a);
				samples[p * 3 + 2] = z;
			}
		}
	}

	cfl_write(argv[optind], DIMS, dims, samples);
	free(samples);

	return 0;
}
