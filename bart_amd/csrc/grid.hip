// grid.hip -- Kaiser-Bessel gridding for gfx950.
//
// Maths of noncart/grid.c:153-253 and :403-512 (tap range, clamp or wrap,
// separable weights from the 1001-entry KB table with linear interpolation),
// with a different execution strategy from gpu_grid.cu:
//   * samples are bin-sorted into tiles of the grid once per trajectory
//     (the plan lives with the operator; BART rebuilds it per call);
//   * adjoint (k-space -> grid): one work-group per (tile chunk, sub-grid,
//     coil chunk) accumulates into an LDS-resident window (tile + halo) with
//     LDS float atomics, then adds the window to the grid with global
//     atomics -- the only global atomics are these dense tile flushes;
//   * forward (grid -> k-space): one thread per sorted sample gathers all
//     taps of all sub-grids, so neighbouring lanes read neighbouring grid
//     lines; no atomics.
#include "nufft_impl.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bam {

// ---------------------------------------------------------------------------
// Kaiser-Bessel table (noncart/grid.c:38-97, 516-519)
// ---------------------------------------------------------------------------

static double bessel_i0(double x)
{
	double q = 0.25 * x * x;
	double term = 1., sum = 1.;

	for (int k = 1; k < 500; k++) {
		term *= q / ((double)k * (double)k);
		sum += term;
		if (term < 1.e-17 * sum)
			break;
	}

	return sum;
}

double calc_beta(float os, float width)
{
	return M_PI * sqrt(pow((double)(width / os) * (os - 0.5), 2.) - 0.8);
}

void kb_table_host(double beta, float* table)
{
	double i0b = bessel_i0(beta);

	for (int i = 0; i < KB_SIZE + 1; i++) {

		double x = (double)i / (double)(KB_SIZE - 1) / 2.;
		table[i] = (fabs(x) >= 0.5) ? 0.f : (float)(bessel_i0(beta * sqrt(1. - pow(2. * x, 2.))) / i0b);
	}
}

// intlookup (noncart/grid.c:126-149): x in [0, 0.5]
__device__ inline float kb_lookup(const float* table, float x)
{
	x *= 2.f;
	float xs = x * (float)(KB_SIZE - 1);
	int index = (int)xs;
	index = min(max(index, 0), KB_SIZE - 1);
	float fpart = xs - (float)index;
	return (1.f - fpart) * table[index] + fpart * table[index + 1];
}

// ---------------------------------------------------------------------------
// Binning plan
// ---------------------------------------------------------------------------

// sample position on the grid, pos = os (k + shift) + n / 2, rounded after
// each operation as in noncart/grid.c:181-187
#pragma clang fp contract(off)
static float host_pos(const GridGeom& g, int s, int d, float k)
{
	float a = k + g.shift[s][d];
	float b = g.os * a;
	return b + (float)(g.n[d] / 2);
}
#pragma clang fp contract(on)

void GridPlan::build(const GridGeom& geom, const float* traj, int tile_override)
{
	g = geom;

	int nd = 0;
	for (int d = 0; d < 3; d++)
		if (g.n[d] > 1)
			nd++;

	// tile size: 2-d 16 x 16, 3-d 8 x 8 x 8, 1-d 64 (overridable)
	int Tdef = (3 == nd) ? 8 : ((2 == nd) ? 16 : 64);

	if (tile_override > 0)
		Tdef = tile_override;

	float maxshift = 0.f;
	for (int s = 0; s < g.S; s++)
		for (int d = 0; d < 3; d++)
			maxshift = std::max(maxshift, fabsf(g.shift[s][d] - g.shift[0][d]));

	g.Wtot = 1;

	for (int d = 0; d < 3; d++) {

		if (g.n[d] > 1) {

			g.T[d] = std::min(Tdef, g.n[d]);
			g.nt[d] = (g.n[d] + g.T[d] - 1) / g.T[d];
			g.lo[d] = (int)ceil(0.5 * g.width + g.os * maxshift) + 1;
			g.hi[d] = (int)ceil(0.5 * g.width) + 1;
			g.W[d] = g.T[d] + g.lo[d] + g.hi[d];

		} else {

			g.T[d] = 1;
			g.nt[d] = 1;
			g.lo[d] = 0;
			g.hi[d] = 0;
			g.W[d] = 1;
		}

		g.Wtot *= g.W[d];
	}

	long ntiles = (long)g.nt[0] * g.nt[1] * g.nt[2];
	std::vector<int> tile(g.M);
	std::vector<long> count(ntiles + 1, 0);

	std::vector<int> wraps(g.M);

	for (long m = 0; m < g.M; m++) {

		int tb[3];
		int code = 0;

		for (int d = 0; d < 3; d++) {

			if (1 == g.n[d]) {
				tb[d] = 0;
				code |= 8 << (4 * d);
				continue;
			}

			float p = host_pos(g, 0, d, traj[3 * m + d]);
			int q = 0;

			if (!std::isfinite(p))
				p = 0.f;

			if (g.periodic) {
				q = (int)floorf(p / (float)g.n[d]);
				q = std::min(std::max(q, -7), 7);
				p = p - (float)g.n[d] * (float)q;
			}

			code |= (q + 8) << (4 * d);

			int b = (int)floorf(p / (float)g.T[d]);
			tb[d] = std::min(std::max(b, 0), g.nt[d] - 1);
		}

		wraps[m] = code;
		tile[m] = tb[0] + g.nt[0] * (tb[1] + g.nt[1] * tb[2]);
		count[tile[m] + 1]++;
	}

	for (long t = 0; t < ntiles; t++)
		count[t + 1] += count[t];

	std::vector<int> hsm(g.M);
	std::vector<float4> hsk(g.M);
	std::vector<long> pos(count.begin(), count.end() - 1);

	for (long m = 0; m < g.M; m++) {

		long i = pos[tile[m]]++;
		hsm[i] = (int)m;
		int code = wraps[m];
		float fcode;
		memcpy(&fcode, &code, sizeof(float));
		hsk[i] = make_float4(traj[3 * m + 0], traj[3 * m + 1], traj[3 * m + 2], fcode);
	}

	// work items: tiles split into chunks of at most CH samples
	const long CH = (3 == nd) ? 512 : 1024;
	std::vector<int4> hw;

	for (long t = 0; t < ntiles; t++)
		for (long b = count[t]; b < count[t + 1]; b += CH)
			hw.push_back(make_int4((int)t, (int)b, (int)std::min(b + CH, count[t + 1]), 0));

	nwork = (int)hw.size();

	// coils per adjoint work-group: keep the LDS window <= 56 KiB
	cc_adj = std::max(1, std::min(g.C, (int)((56 * 1024) / ((long)g.Wtot * 8))));
	cc_fwd = std::min(g.C, 16);

	BAM_CHECK((long)g.Wtot * 8 <= 56 * 1024, "gridding window too large (%d points)", g.Wtot);
	BAM_CHECK(g.M < (1l << 31), "too many samples");

	hipStream_t st = 0;
	sm.alloc(std::max<long>(g.M, 1));
	sk.alloc(std::max<long>(g.M, 1));
	work.alloc(std::max<int>(nwork, 1));

	if (g.M > 0) {
		BAM_HIP(hipMemcpyAsync(sm.ptr, hsm.data(), sizeof(int) * g.M, hipMemcpyHostToDevice, st));
		BAM_HIP(hipMemcpyAsync(sk.ptr, hsk.data(), sizeof(float4) * g.M, hipMemcpyHostToDevice, st));
	}
	if (nwork > 0)
		BAM_HIP(hipMemcpyAsync(work.ptr, hw.data(), sizeof(int4) * nwork, hipMemcpyHostToDevice, st));

	BAM_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Taps of one sample along one dim (noncart/grid.c:415-457)
// ---------------------------------------------------------------------------

struct Taps {
	int n;
	int first;      // first tap index (unwrapped, after clamping)
	float w[MAX_TAPS];
};

__device__ inline void taps_dim(Taps& t, float pos, int dim, bool periodic, float width, const float* table)
{
	int sti = (int)ceil((double)pos - 0.5 * (double)width);
	int eni = (int)floor((double)pos + 0.5 * (double)width);

	if (!periodic) {
		sti = max(sti, 0);
		eni = min(eni, dim - 1);
	}

	int cnt = eni - sti + 1;
	cnt = min(cnt, MAX_TAPS);
	t.n = max(cnt, 0);
	t.first = sti;

#pragma unroll
	for (int j = 0; j < MAX_TAPS; j++) {
		if (j < t.n) {
			float frac = fabsf((float)(sti + j) - pos);
			t.w[j] = kb_lookup(table, frac / width);
		}
	}
}

__device__ inline float sample_pos(const GridGeom& g, int s, int d, float k)
{
	return __fadd_rn(__fmul_rn(g.os, __fadd_rn(k, g.shift[s][d])), (float)(g.n[d] / 2));
}

// ---------------------------------------------------------------------------
// Adjoint gridding (k-space -> grid), LDS-tiled.
// grid layout: [S][C][N]; ksp: [C][M]; launch (nwork, S, ceil(C / cc)).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_grid_adj(GridGeom g, int cc, const float* __restrict__ kbt,
						  const int* __restrict__ sm, const float4* __restrict__ sk,
						  const int4* __restrict__ work, const float2* __restrict__ ksp,
						  const float2* __restrict__ weights, float2* __restrict__ grid)
{
	extern __shared__ float smem[];
	float* table = smem;                             // KB_SIZE + 1 (padded to 1004)
	float2* win = (float2*)(smem + 1004);            // cc * Wtot

	const int4 wk = work[blockIdx.x];
	const int s = blockIdx.y;
	const int c0 = blockIdx.z * cc;
	const int ncc = min(cc, g.C - c0);

	for (int i = threadIdx.x; i < KB_SIZE + 1; i += blockDim.x)
		table[i] = kbt[i];

	for (int i = threadIdx.x; i < ncc * g.Wtot; i += blockDim.x)
		win[i] = make_float2(0.f, 0.f);

	int tb[3];
	tb[0] = wk.x % g.nt[0];
	tb[1] = (wk.x / g.nt[0]) % g.nt[1];
	tb[2] = wk.x / (g.nt[0] * g.nt[1]);

	int org[3];
	for (int d = 0; d < 3; d++)
		org[d] = tb[d] * g.T[d] - g.lo[d];

	__syncthreads();

	for (int i = wk.y + threadIdx.x; i < wk.z; i += blockDim.x) {

		const int m = sm[i];
		const float4 kc = sk[i];
		const float kk[3] = { kc.x, kc.y, kc.z };

		Taps t[3];
		int base[3];
		bool any = true;

		for (int d = 0; d < 3; d++) {

			if (1 == g.n[d]) {
				t[d].n = 1;
				t[d].first = 0;
				t[d].w[0] = 1.f;
				base[d] = 0;
				continue;
			}

			float pos = sample_pos(g, s, d, kk[d]);
			taps_dim(t[d], pos, g.n[d], g.periodic, g.width, table);

			int q = ((__float_as_int(kc.w) >> (4 * d)) & 15) - 8;

			base[d] = t[d].first - q * g.n[d] - org[d];
			any = any && (t[d].n > 0);
		}

		if (!any)
			continue;

		float2 w = make_float2(1.f, 0.f);

		if (nullptr != weights)
			w = weights[m];

		for (int c = 0; c < ncc; c++) {

			float2 v = cmulc(ksp[(long)(c0 + c) * g.M + m], w);

			if (0.f == v.x && 0.f == v.y)
				continue;

			float2* wc = win + c * g.Wtot;

			for (int iz = 0; iz < t[2].n; iz++) {

				int lz = base[2] + iz;
				if (lz < 0 || lz >= g.W[2])
					continue;
				float dz = 1.f * t[2].w[iz];

				for (int iy = 0; iy < t[1].n; iy++) {

					int ly = base[1] + iy;
					if (ly < 0 || ly >= g.W[1])
						continue;
					float dy = dz * t[1].w[iy];
					float2* row = wc + g.W[0] * (ly + g.W[1] * lz);

					for (int ix = 0; ix < t[0].n; ix++) {

						int lx = base[0] + ix;
						if (lx < 0 || lx >= g.W[0])
							continue;
						float dd = dy * t[0].w[ix];
						atomicAdd(&row[lx].x, v.x * dd);
						atomicAdd(&row[lx].y, v.y * dd);
					}
				}
			}
		}
	}

	__syncthreads();

	// flush the window: wrap (periodic) or drop (non-periodic) out-of-range points
	for (int j = threadIdx.x; j < ncc * g.Wtot; j += blockDim.x) {

		float2 v = win[j];

		if (0.f == v.x && 0.f == v.y)
			continue;

		int c = j / g.Wtot;
		int r = j - c * g.Wtot;
		int l[3];
		l[0] = r % g.W[0];
		l[1] = (r / g.W[0]) % g.W[1];
		l[2] = r / (g.W[0] * g.W[1]);

		long idx = 0;
		long str = 1;
		bool ok = true;

		for (int d = 0; d < 3; d++) {

			int gi = org[d] + l[d];

			if (g.periodic) {
				gi %= g.n[d];
				if (gi < 0)
					gi += g.n[d];
			} else if (gi < 0 || gi >= g.n[d]) {
				ok = false;
			}

			idx += gi * str;
			str *= g.n[d];
		}

		if (!ok)
			continue;

		float2* dst = grid + ((long)s * g.C + c0 + c) * g.N + idx;
		unsafeAtomicAdd(&dst->x, v.x);
		unsafeAtomicAdd(&dst->y, v.y);
	}
}

void grid_adjoint(const GridPlan& gp, const float* kb, float2* grid, const float2* ksp, const float2* weights,
		  hipStream_t st)
{
	if (0 == gp.nwork || 0 == gp.g.C)
		return;

	dim3 blocks(gp.nwork, gp.g.S, ceil_div(gp.g.C, gp.cc_adj));
	size_t lds = 1004 * sizeof(float) + (size_t)gp.cc_adj * gp.g.Wtot * sizeof(float2);

	hipLaunchKernelGGL(k_grid_adj, blocks, dim3(256), lds, st, gp.g, gp.cc_adj, kb, gp.sm.ptr, gp.sk.ptr,
			   gp.work.ptr, ksp, weights, grid);
	BAM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Forward gridding (grid -> k-space): one thread per sorted sample and coil
// chunk; sums the sub-grids in order s = 0 .. S-1 (noncart/nufft.c:1444-1445)
// and multiplies by the weights (:1456-1457).
// ---------------------------------------------------------------------------

template <int CC>
__global__ __launch_bounds__(256) void k_grid_fwd(GridGeom g, const float* __restrict__ kbt,
						  const int* __restrict__ sm, const float4* __restrict__ sk,
						  const float2* __restrict__ grid, const float2* __restrict__ weights,
						  float2* __restrict__ ksp)
{
	__shared__ float table[KB_SIZE + 1];

	for (int i = threadIdx.x; i < KB_SIZE + 1; i += blockDim.x)
		table[i] = kbt[i];

	__syncthreads();

	const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;

	if (i >= g.M)
		return;

	const int c0 = blockIdx.y * CC;
	const int ncc = min(CC, g.C - c0);
	const int m = sm[i];
	const float4 kc = sk[i];
	const float kk[3] = { kc.x, kc.y, kc.z };

	float2 acc[CC];

#pragma unroll
	for (int c = 0; c < CC; c++)
		acc[c] = make_float2(0.f, 0.f);

	for (int s = 0; s < g.S; s++) {

		Taps t[3];
		bool any = true;

		for (int d = 0; d < 3; d++) {

			if (1 == g.n[d]) {
				t[d].n = 1;
				t[d].first = 0;
				t[d].w[0] = 1.f;
				continue;
			}

			float pos = sample_pos(g, s, d, kk[d]);
			taps_dim(t[d], pos, g.n[d], g.periodic, g.width, table);
			any = any && (t[d].n > 0);
		}

		if (!any)
			continue;

		float2 val[CC];

#pragma unroll
		for (int c = 0; c < CC; c++)
			val[c] = make_float2(0.f, 0.f);

		const float2* gs = grid + ((long)s * g.C + c0) * g.N;

		for (int iz = 0; iz < t[2].n; iz++) {

			int z = t[2].first + iz;
			if (g.periodic) {
				z %= g.n[2];
				if (z < 0)
					z += g.n[2];
			}
			float dz = 1.f * t[2].w[iz];

			for (int iy = 0; iy < t[1].n; iy++) {

				int y = t[1].first + iy;
				if (g.periodic) {
					y %= g.n[1];
					if (y < 0)
						y += g.n[1];
				}
				float dy = dz * t[1].w[iy];
				const float2* row = gs + (long)g.n[0] * (y + (long)g.n[1] * z);

				for (int ix = 0; ix < t[0].n; ix++) {

					int x = t[0].first + ix;
					if (g.periodic) {
						x %= g.n[0];
						if (x < 0)
							x += g.n[0];
					}
					float dd = dy * t[0].w[ix];

#pragma unroll
					for (int c = 0; c < CC; c++) {
						if (c < ncc) {
							float2 v = row[(long)c * g.N + x];
							val[c].x += v.x * dd;
							val[c].y += v.y * dd;
						}
					}
				}
			}
		}

#pragma unroll
		for (int c = 0; c < CC; c++)
			acc[c] = cadd(acc[c], val[c]);
	}

	float2 w = make_float2(1.f, 0.f);

	if (nullptr != weights)
		w = weights[m];

#pragma unroll
	for (int c = 0; c < CC; c++)
		if (c < ncc)
			ksp[(long)(c0 + c) * g.M + m] = (nullptr != weights) ? cmul(acc[c], w) : acc[c];
}

void grid_forward(const GridPlan& gp, const float* kb, float2* ksp, const float2* grid, const float2* weights,
		  hipStream_t st)
{
	if (0 == gp.g.M || 0 == gp.g.C)
		return;

	int cc = gp.cc_fwd;

	if (cc <= 4)
		hipLaunchKernelGGL((k_grid_fwd<4>), dim3(ceil_div(gp.g.M, 256), ceil_div(gp.g.C, 4)), dim3(256), 0, st, gp.g, kb, gp.sm.ptr, gp.sk.ptr, grid, weights, ksp);
	else if (cc <= 8)
		hipLaunchKernelGGL((k_grid_fwd<8>), dim3(ceil_div(gp.g.M, 256), ceil_div(gp.g.C, 8)), dim3(256), 0, st, gp.g, kb, gp.sm.ptr, gp.sk.ptr, grid, weights, ksp);
	else
		hipLaunchKernelGGL((k_grid_fwd<16>), dim3(ceil_div(gp.g.M, 256), ceil_div(gp.g.C, 16)), dim3(256), 0, st, gp.g, kb, gp.sm.ptr, gp.sk.ptr, grid, weights, ksp);

	BAM_HIP(hipGetLastError());
}

}  // namespace bam
