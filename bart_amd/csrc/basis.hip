// basis.hip -- subspace (temporal basis) NUFFT: the contraction of the
// coefficient k-space with the basis, its adjoint, the coil maps of the
// coefficient images, and the K x K Toeplitz PSF product
// (noncart/nufft.c:470-505, 1450-1453, 1483-1487, 1546-1625;
// sense/modelnc.c:28-66).
//
// Layouts (device, complex float):
//   basis B       [K][T]              B[t + T k] (BART dims [1 1 1 1 1 T K])
//   frame k-space [T][C][Mf]          m + Mf (c + C t) (dims [1 R S C 1 T])
//   coeff k-space [K][C][T][Mf]       the plan's channels c + C k over its
//                                     M' = T Mf samples m + Mf t
//   coeff images  [K][C][N]           (dims [X Y Z C 1 1 K])
//   PSF           [K K][S][N]         entry k K + k' multiplies coefficient
//                                     k' into coefficient k
#include "nufft_impl.hpp"

namespace bam {

static unsigned blocks_for(long n)
{
	return (unsigned)std::min<long>(std::max<long>((n + 255) / 256, 1), 8192);
}

__device__ inline bool skipped(const int* skip) { return nullptr != skip && 0 != *skip; }

// dst[m + Mf (c + C t)] = sum_k B[t + T k] src[((c + C k) T + t) Mf + m]
__global__ void k_basis_contract(float2* __restrict__ dst, const float2* __restrict__ src,
				 const float2* __restrict__ B, int C, int K, int T, long Mf, const int* skip)
{
	if (skipped(skip))
		return;

	const long total = Mf * C * T;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long m = e % Mf;
		const long ct = e / Mf;
		const int c = (int)(ct % C);
		const int t = (int)(ct / C);
		float2 acc = make_float2(0.f, 0.f);

		for (int k = 0; k < K; k++)
			acc = cadd(acc, cmul(src[((long)(c + C * k) * T + t) * Mf + m], B[t + (long)T * k]));

		dst[e] = acc;
	}
}

// dst[((c + C k) T + t) Mf + m] = src[m + Mf (c + C t)] conj(B[t + T k])
__global__ void k_basis_expand(float2* __restrict__ dst, const float2* __restrict__ src,
			       const float2* __restrict__ B, int C, int K, int T, long Mf, const int* skip)
{
	if (skipped(skip))
		return;

	const long total = Mf * T * C * K;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long m = e % Mf;
		const long r = e / Mf;
		const int t = (int)(r % T);
		const long ck = r / T;
		const int c = (int)(ck % C);
		const int k = (int)(ck / C);
		dst[e] = cmulc(src[m + Mf * (c + (long)C * t)], B[t + (long)T * k]);
	}
}

// dst[(c + C k) N + i] = sens[c N + i] x[k N + i]
__global__ void k_maps_expand(float2* __restrict__ dst, const float2* __restrict__ sens,
			      const float2* __restrict__ x, int C, int K, long N, const int* skip)
{
	if (skipped(skip))
		return;

	const long total = N * C * K;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long i = e % N;
		const long ck = e / N;
		const int c = (int)(ck % C);
		const int k = (int)(ck / C);
		dst[e] = cmul(sens[(long)c * N + i], x[(long)k * N + i]);
	}
}

// x[k N + i] = sum_c conj(sens[c N + i]) src[(c + C k) N + i], c in order
__global__ void k_maps_reduce(float2* __restrict__ x, const float2* __restrict__ sens,
			      const float2* __restrict__ src, int C, int K, long N, const int* skip)
{
	if (skipped(skip))
		return;

	const long total = N * K;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long i = e % N;
		const int k = (int)(e / N);
		float2 acc = make_float2(0.f, 0.f);

		for (int c = 0; c < C; c++)
			acc = cadd(acc, cmulc(src[((long)c + (long)C * k) * N + i], sens[(long)c * N + i]));

		x[e] = acc;
	}
}

// grid[s][c + C k][i] = sum_k' psf[k K + k'][s][i] grid[s][c + C k'][i]
// (md_ztenmul over the coefficient dims, noncart/nufft.c:1594-1602); the
// stack holds S x (C K) volumes of N points
template <int KK>  // KK >= K: register arrays of fixed size
__global__ void k_basis_psf(float2* __restrict__ grid, const float2* __restrict__ psf, int S, int C, int K, long N,
			    const int* skip)
{
	if (skipped(skip))
		return;

	const long total = (long)S * C * N;
	const long CK = (long)C * K;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long i = e % N;
		const long sc = e / N;
		const int c = (int)(sc % C);
		const int s = (int)(sc / C);
		float2 v[KK];

#pragma unroll
		for (int k = 0; k < KK; k++)
			v[k] = (k < K) ? grid[((long)s * CK + c + (long)C * k) * N + i] : make_float2(0.f, 0.f);

#pragma unroll
		for (int k = 0; k < KK; k++) {

			if (k >= K)
				break;

			float2 acc = make_float2(0.f, 0.f);

#pragma unroll
			for (int q = 0; q < KK; q++)
				if (q < K)
					acc = cadd(acc, cmul(psf[(((long)k * K + q) * S + s) * N + i], v[q]));

			grid[((long)s * CK + c + (long)C * k) * N + i] = acc;
		}
	}
}

// the k-space of the K x K PSF channels (compute_square_basis,
// noncart/nufft.c:470-505, and the conj of the adjoint's basis product):
// v[(k K + k') M' + m'] = conj(B[t + T k]) B[t + T k'], t = m' / Mf
__global__ void k_basis_psf_vals(float2* __restrict__ v, const float2* __restrict__ B, int K, int T, long Mf)
{
	const long Mp = Mf * T;
	const long total = Mp * K * K;

	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		const long m = e % Mp;
		const long j = e / Mp;
		const int k = (int)(j / K);
		const int q = (int)(j % K);
		const int t = (int)(m / Mf);
		v[e] = cmulc(B[t + (long)T * q], B[t + (long)T * k]);
	}
}

void basis_contract(float2* dst, const float2* src, const float2* B, int C, int K, int T, long Mf, hipStream_t st,
		    const int* skip)
{
	hipLaunchKernelGGL(k_basis_contract, dim3(blocks_for(Mf * C * T)), dim3(256), 0, st, dst, src, B, C, K, T, Mf,
			   skip);
	BAM_HIP(hipGetLastError());
}

void basis_expand(float2* dst, const float2* src, const float2* B, int C, int K, int T, long Mf, hipStream_t st,
		  const int* skip)
{
	hipLaunchKernelGGL(k_basis_expand, dim3(blocks_for(Mf * T * C * K)), dim3(256), 0, st, dst, src, B, C, K, T,
			   Mf, skip);
	BAM_HIP(hipGetLastError());
}

void maps_expand(float2* dst, const float2* sens, const float2* x, int C, int K, long N, hipStream_t st,
		 const int* skip)
{
	hipLaunchKernelGGL(k_maps_expand, dim3(blocks_for(N * C * K)), dim3(256), 0, st, dst, sens, x, C, K, N, skip);
	BAM_HIP(hipGetLastError());
}

void maps_reduce(float2* x, const float2* sens, const float2* src, int C, int K, long N, hipStream_t st,
		 const int* skip)
{
	hipLaunchKernelGGL(k_maps_reduce, dim3(blocks_for(N * K)), dim3(256), 0, st, x, sens, src, C, K, N, skip);
	BAM_HIP(hipGetLastError());
}

void basis_psf_apply(float2* grid, const float2* psf, int S, int C, int K, long N, hipStream_t st, const int* skip)
{
	BAM_CHECK(K >= 1 && K <= BASIS_MAX, "basis: %d coefficients (at most %d)", K, BASIS_MAX);
	auto kern = (K <= 2) ? k_basis_psf<2> : (K <= 4) ? k_basis_psf<4> : (K <= 8) ? k_basis_psf<8> : k_basis_psf<BASIS_MAX>;
	hipLaunchKernelGGL(kern, dim3(blocks_for((long)S * C * N)), dim3(256), 0, st, grid, psf, S, C, K, N,
			   skip);
	BAM_HIP(hipGetLastError());
}

void basis_psf_vals(float2* v, const float2* B, int K, int T, long Mf, hipStream_t st)
{
	hipLaunchKernelGGL(k_basis_psf_vals, dim3(blocks_for(Mf * T * K * K)), dim3(256), 0, st, v, B, K, T, Mf);
	BAM_HIP(hipGetLastError());
}

}  // namespace bam
