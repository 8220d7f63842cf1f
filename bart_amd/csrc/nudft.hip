// nudft.hip -- direct non-uniform DFT (noncart/nudft.c:69-190), the
// operator behind "nufft -s".
//
//   forward:  ksp[c][m] = s sum_x img[c][x] exp(i phi(m, x))
//   adjoint:  img[c][x] = s sum_m ksp[c][m] exp(-i phi(m, x))
//   phi(m, x) = -2 pi sum_d k_d(m) (x_d - n_d / 2) / n_d
//
// (get_coord negates the coordinates, linear_phase / centered_gradient
// centre them at n_d / 2, num/filter.c:241-265.)  The scale s is
// 1 / sqrt(size of the whole coil-image array), as nudft_forward2 scales by
// md_calc_size(N, idims) with the coil dim included.  The phase is formed in
// FP64 and reduced to [-pi, pi] before the FP32 sine and cosine; the sums
// are FP32.  No B0 term.
#include "nudft.hpp"
#include "prof.hpp"

#include <cmath>
#include <vector>

namespace bam {

constexpr int NUDFT_THREADS = 256;
constexpr int NUDFT_TILE = 512;

__device__ inline float2 expi_reduced(double ph)
{
	const double two_pi = 6.283185307179586476925286766559;
	const double r = ph - two_pi * rint(ph * (1.0 / two_pi));
	float s, c;
	sincosf((float)r, &s, &c);
	return make_float2(c, s);
}

// thread per sample, image tiles staged in LDS; grid (ceil(M / 256), C)
__global__ __launch_bounds__(NUDFT_THREADS) void k_nudft_fwd(NudftGeom g, const double4* __restrict__ grad,
							     const float2* __restrict__ img, float2* __restrict__ ksp)
{
	__shared__ float2 tile[NUDFT_TILE];

	const long m = (long)blockIdx.x * NUDFT_THREADS + threadIdx.x;
	const int c = blockIdx.y;
	const bool active = m < g.M;
	const double4 gm = active ? grad[m] : make_double4(0., 0., 0., 0.);
	const float2* src = img + (long)c * g.N;

	float2 acc = make_float2(0.f, 0.f);

	for (long x0 = 0; x0 < g.N; x0 += NUDFT_TILE) {

		const int nt = (int)min((long)NUDFT_TILE, g.N - x0);

		__syncthreads();

		for (int i = threadIdx.x; i < nt; i += NUDFT_THREADS)
			tile[i] = src[x0 + i];

		__syncthreads();

		if (!active)
			continue;

		for (int i = 0; i < nt; i++) {

			const long x = x0 + i;
			const long px = x % g.n[0];
			const long py = (x / g.n[0]) % g.n[1];
			const long pz = x / ((long)g.n[0] * g.n[1]);
			const double ph = gm.x * ((double)px - g.h[0]) + gm.y * ((double)py - g.h[1]) +
					  gm.z * ((double)pz - g.h[2]);
			const float2 e = expi_reduced(ph);
			acc = cadd(acc, cmul(tile[i], e));
		}
	}

	if (active)
		ksp[(long)c * g.M + m] = cscale(acc, g.scale);
}

// thread per pixel, sample tiles staged in LDS; grid (ceil(N / 256), C)
__global__ __launch_bounds__(NUDFT_THREADS) void k_nudft_adj(NudftGeom g, const double4* __restrict__ grad,
							     const float2* __restrict__ ksp, float2* __restrict__ img)
{
	__shared__ double4 gt[NUDFT_TILE];
	__shared__ float2 vt[NUDFT_TILE];

	const long x = (long)blockIdx.x * NUDFT_THREADS + threadIdx.x;
	const int c = blockIdx.y;
	const bool active = x < g.N;
	const double fx = active ? (double)(x % g.n[0]) - g.h[0] : 0.;
	const double fy = active ? (double)((x / g.n[0]) % g.n[1]) - g.h[1] : 0.;
	const double fz = active ? (double)(x / ((long)g.n[0] * g.n[1])) - g.h[2] : 0.;
	const float2* src = ksp + (long)c * g.M;

	float2 acc = make_float2(0.f, 0.f);

	for (long m0 = 0; m0 < g.M; m0 += NUDFT_TILE) {

		const int nt = (int)min((long)NUDFT_TILE, g.M - m0);

		__syncthreads();

		for (int i = threadIdx.x; i < nt; i += NUDFT_THREADS) {
			gt[i] = grad[m0 + i];
			vt[i] = src[m0 + i];
		}

		__syncthreads();

		if (!active)
			continue;

		for (int i = 0; i < nt; i++) {
			const double4 gm = gt[i];
			const double ph = gm.x * fx + gm.y * fy + gm.z * fz;
			acc = cadd(acc, cmulc(vt[i], expi_reduced(ph)));
		}
	}

	if (active)
		img[(long)c * g.N + x] = cscale(acc, g.scale);
}

void Nudft::create(const int dims[3], int ncoils, long nsamples, const float* traj)
{
	for (int d = 0; d < 3; d++) {
		g.n[d] = dims[d];
		g.h[d] = (double)dims[d] / 2.;
	}

	g.N = (long)dims[0] * dims[1] * dims[2];
	g.C = ncoils;
	g.M = nsamples;
	g.scale = 1.f / sqrtf((float)(g.N * g.C));

	// per-sample phase gradients -2 pi k_d / n_d (dims of size 1: the
	// coordinate still enters with x_d - 1/2, as in centered_gradient)
	std::vector<double4> h(std::max<long>(nsamples, 1));

	for (long m = 0; m < nsamples; m++) {
		double gd[3];
		for (int d = 0; d < 3; d++)
			gd[d] = -2. * M_PI * (double)traj[3 * m + d] / (double)dims[d];
		h[m] = make_double4(gd[0], gd[1], gd[2], 0.);
	}

	grad.alloc(std::max<long>(nsamples, 1));
	BAM_HIP(hipMemcpy(grad.ptr, h.data(), sizeof(double4) * std::max<long>(nsamples, 1), hipMemcpyHostToDevice));
}

void Nudft::forward(float2* ksp, const float2* img)
{
	if (0 == g.M || 0 == g.C)
		return;

	ProfScope prof("nudft_forward", stream);
	hipLaunchKernelGGL(k_nudft_fwd, dim3((unsigned)ceil_div(g.M, NUDFT_THREADS), g.C), dim3(NUDFT_THREADS), 0,
			   stream, g, grad.ptr, img, ksp);
	BAM_HIP(hipGetLastError());
}

void Nudft::adjoint(float2* img, const float2* ksp)
{
	if (0 == g.N || 0 == g.C)
		return;

	ProfScope prof("nudft_adjoint", stream);
	hipLaunchKernelGGL(k_nudft_adj, dim3((unsigned)ceil_div(g.N, NUDFT_THREADS), g.C), dim3(NUDFT_THREADS), 0,
			   stream, g, grad.ptr, ksp, img);
	BAM_HIP(hipGetLastError());
}

}  // namespace bam
