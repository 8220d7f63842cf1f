// cg.hip -- conjugate gradients on (A^H A + lambda I) x = b, gfx950.
//
// Arithmetic of iter/italgos.c:497-580 (conjgrad) with the stopping rule of
// iter/iter2.c:227-270 (stop when ||r||^2 <= (tol ||b||)^2; tol = 0 runs
// maxiter iterations) and the FP64 dot/norm accumulation of
// num/vecops.c:89-124.  Differences in execution, not in maths:
//   * the scalars (rsold, alpha, beta, stop flags) stay on the device; the
//     host does not read anything until the solve ends (BART's GPU path copies
//     two scalars to the host every iteration, num/gpukrnls.cu:1847-1858);
//   * reductions are two-stage and deterministic (per-block FP64 partials,
//     then a fixed-order sum);
//   * maxiter iterations -- normal operator and vector updates -- are captured
//     once into a hipGraph and replayed; after convergence every kernel sees
//     the device stop flag and returns immediately.
#include "cg.hpp"
#include "prof.hpp"

#include <cmath>

namespace bam {

constexpr int CG_THREADS = 256;
constexpr int CG_BLOCKS = 512;

__device__ inline double block_sum(double v, double* sh)
{
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_down(v, off, 64);

	const int lane = threadIdx.x & 63;
	const int wid = threadIdx.x >> 6;

	if (0 == lane)
		sh[wid] = v;

	__syncthreads();

	double r = 0.;

	if (0 == threadIdx.x)
		for (int w = 0; w < (int)(blockDim.x >> 6); w++)
			r += sh[w];

	return r;  // valid in thread 0
}

// deterministic sum of the partials (thread 0 of a single block)
__device__ inline double sum_partials(const double* part, int n, double* sh)
{
	double v = 0.;

	for (int i = threadIdx.x; i < n; i += blockDim.x)
		v += part[i];

	return block_sum(v, sh);
}

// Ap += lambda p; partial sums of p . Ap (float products, FP64 accumulation)
__global__ __launch_bounds__(CG_THREADS) void k_cg_apdot(long n, float lambda, const float* __restrict__ p,
							 float* __restrict__ Ap, double* __restrict__ part,
							 const CgState* st)
{
	__shared__ double sh[4];

	if (0 != st->done)
		return;

	double acc = 0.;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {

		float pv = p[i];
		float a = Ap[i] + lambda * pv;
		Ap[i] = a;
		acc += (double)(pv * a);
	}

	double s = block_sum(acc, sh);

	if (0 == threadIdx.x)
		part[blockIdx.x] = s;
}

__global__ __launch_bounds__(CG_THREADS) void k_cg_alpha(int nparts, const double* __restrict__ part, CgState* st)
{
	__shared__ double sh[4];

	if (0 != st->done)
		return;

	double s = sum_partials(part, nparts, sh);

	if (0 == threadIdx.x) {

		float pAp = (float)s;

		if (0.f == pAp) {
			st->done = 3;     // breakdown: x unchanged (italgos.c:544-545)
		} else {
			st->alpha = st->rsold / pAp;
		}
	}
}

// x += alpha p; r -= alpha Ap; partial sums of r . r
__global__ __launch_bounds__(CG_THREADS) void k_cg_xr(long n, float* __restrict__ x, float* __restrict__ r,
						      const float* __restrict__ p, const float* __restrict__ Ap,
						      double* __restrict__ part, const CgState* st)
{
	__shared__ double sh[4];

	if (0 != st->done)
		return;

	const float alpha = st->alpha;
	double acc = 0.;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {

		x[i] += alpha * p[i];
		float rv = r[i] + (-alpha) * Ap[i];
		r[i] = rv;
		acc += (double)(rv * rv);
	}

	double s = block_sum(acc, sh);

	if (0 == threadIdx.x)
		part[blockIdx.x] = s;
}

__global__ __launch_bounds__(CG_THREADS) void k_cg_beta(int nparts, const double* __restrict__ part, CgState* st,
							float* __restrict__ trace)
{
	__shared__ double sh[4];

	if (0 != st->done)
		return;

	double s = sum_partials(part, nparts, sh);

	if (0 == threadIdx.x) {

		double nr = sqrt(s);
		float rsnew = (float)(nr * nr);
		int it = st->iter;

		st->beta = rsnew / st->rsold;
		st->rsold = rsnew;
		st->rsnew = rsnew;
		st->iter = it + 1;

		if (nullptr != trace)
			trace[it] = sqrtf(rsnew);

		if (rsnew <= st->eps2)
			st->done = 2;  // converged
		else if (st->iter >= st->maxiter)
			st->done = 1;  // iteration budget used
	}
}

// p = beta p + r
__global__ __launch_bounds__(CG_THREADS) void k_cg_p(long n, float* __restrict__ p, const float* __restrict__ r,
						     const CgState* st)
{
	if (0 != st->done)
		return;

	const float beta = st->beta;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
		p[i] = beta * p[i] + r[i];
}

// r = b - (A x + lambda x) [with ax = A x already in r]; p = r; ||r||^2 and ||b||^2 partials
__global__ __launch_bounds__(CG_THREADS) void k_cg_init(long n, float lambda, const float* __restrict__ x,
							const float* __restrict__ b, float* __restrict__ r,
							float* __restrict__ p, double* __restrict__ part,
							double* __restrict__ part_b, bool cold)
{
	__shared__ double sh[4];
	double acc = 0., accb = 0.;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {

		float bv = b[i];
		float rv;

		if (cold) {
			rv = bv;
		} else {
			float ax = r[i] + lambda * x[i];
			rv = -1.f * ax + bv;
		}

		r[i] = rv;
		p[i] = rv;
		acc += (double)(rv * rv);
		accb += (double)(bv * bv);
	}

	double s = block_sum(acc, sh);
	__syncthreads();
	double sb = block_sum(accb, sh);

	if (0 == threadIdx.x) {
		part[blockIdx.x] = s;
		part_b[blockIdx.x] = sb;
	}
}

__global__ void k_cg_init_fin(int nparts, const double* part, const double* part_b, float tol, int maxiter,
			      CgState* st)
{
	__shared__ double sh[4];

	double s = sum_partials(part, nparts, sh);
	__syncthreads();
	double sb = sum_partials(part_b, nparts, sh);

	if (0 == threadIdx.x) {

		double nb = sqrt(sb);
		float eps = (float)nb * tol;
		double nr = sqrt(s);
		float rs = (float)(nr * nr);

		st->bnorm = (float)nb;
		st->r0norm = (float)nr;
		st->rsold = rs;
		st->rsnew = rs;
		st->eps2 = (float)pow((double)eps, 2.);
		st->iter = 0;
		st->maxiter = maxiter;
		st->alpha = 0.f;
		st->beta = 0.f;
		// checkeps (iter2.c:65-80) and the early out of italgos.c:526-530
		st->done = (0. == nb || 0.f == rs || maxiter <= 0) ? 4 : 0;
	}
}

// ---------------------------------------------------------------------------

CgSolver::CgSolver(long nfloats, hipStream_t s) : n(nfloats), stream(s)
{
	x.alloc(n);
	b.alloc(n);
	r.alloc(n);
	p.alloc(n);
	Ap.alloc(n);
	part.alloc(CG_BLOCKS);
	part_b.alloc(CG_BLOCKS);
	st.alloc(1);
	trace.alloc(1);
	nblocks = (int)std::min<long>(CG_BLOCKS, std::max<long>(1, (n + CG_THREADS - 1) / CG_THREADS));
	BAM_HIP(hipHostMalloc((void**)&host_st, sizeof(CgState), hipHostMallocDefault));
}

CgSolver::~CgSolver()
{
	drop_graph();
	if (nullptr != host_st)
		(void)hipHostFree(host_st);
}

void CgSolver::drop_graph()
{
	if (nullptr != gexec)
		(void)hipGraphExecDestroy(gexec);
	if (nullptr != graph)
		(void)hipGraphDestroy(graph);
	gexec = nullptr;
	graph = nullptr;
	graph_iters = -1;
}

void CgSolver::enqueue_iteration(const NormalFn& normal, float lambda)
{
	normal((float2*)Ap.ptr, (const float2*)p.ptr, &st.ptr->done);
	ProfScope prof("cg_vector_ops", stream);
	hipLaunchKernelGGL(k_cg_apdot, dim3(nblocks), dim3(CG_THREADS), 0, stream, n, lambda, p.ptr, Ap.ptr, part.ptr, st.ptr);
	hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(CG_THREADS), 0, stream, nblocks, part.ptr, st.ptr);
	hipLaunchKernelGGL(k_cg_xr, dim3(nblocks), dim3(CG_THREADS), 0, stream, n, x.ptr, r.ptr, p.ptr, Ap.ptr, part.ptr, st.ptr);
	hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(CG_THREADS), 0, stream, nblocks, part.ptr, st.ptr, trace.ptr);
	hipLaunchKernelGGL(k_cg_p, dim3(nblocks), dim3(CG_THREADS), 0, stream, n, p.ptr, r.ptr, st.ptr);
	BAM_HIP(hipGetLastError());
}

CgResult CgSolver::solve(const NormalFn& normal, float lambda, int maxiter, float tol, float2* x_io,
			 const float2* b_in, bool warm, bool fallback, bool use_graph)
{
	CgResult res;

	if ((long)trace.n < std::max(1, maxiter)) {
		trace.alloc(std::max(1, maxiter));
		drop_graph();
	}

	hipEvent_t e0, e1;
	BAM_HIP(hipEventCreate(&e0));
	BAM_HIP(hipEventCreate(&e1));
	BAM_HIP(hipEventRecord(e0, stream));

	BAM_HIP(hipMemcpyAsync(b.ptr, b_in, sizeof(float) * n, hipMemcpyDeviceToDevice, stream));

	bool cold = !warm;

	if (warm) {

		BAM_HIP(hipMemcpyAsync(x.ptr, x_io, sizeof(float) * n, hipMemcpyDeviceToDevice, stream));
		normal((float2*)r.ptr, (const float2*)x.ptr, nullptr);

	} else {

		fill((float2*)x.ptr, n / 2, make_float2(0.f, 0.f), stream);
	}

	hipLaunchKernelGGL(k_cg_init, dim3(nblocks), dim3(CG_THREADS), 0, stream, n, lambda, x.ptr, b.ptr, r.ptr, p.ptr,
			   part.ptr, part_b.ptr, cold);
	hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(CG_THREADS), 0, stream, nblocks, part.ptr, part_b.ptr, tol, maxiter,
			   st.ptr);
	BAM_HIP(hipGetLastError());

	if (warm && fallback) {

		// zero start when the warm start's residual is worse than ||b||
		BAM_HIP(hipMemcpyAsync(host_st, st.ptr, sizeof(CgState), hipMemcpyDeviceToHost, stream));
		BAM_HIP(hipStreamSynchronize(stream));

		if (host_st->r0norm > host_st->bnorm) {

			res.fell_back = true;
			fill((float2*)x.ptr, n / 2, make_float2(0.f, 0.f), stream);
			hipLaunchKernelGGL(k_cg_init, dim3(nblocks), dim3(CG_THREADS), 0, stream, n, lambda, x.ptr, b.ptr,
					   r.ptr, p.ptr, part.ptr, part_b.ptr, true);
			hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(CG_THREADS), 0, stream, nblocks, part.ptr, part_b.ptr,
					   tol, maxiter, st.ptr);
			BAM_HIP(hipGetLastError());
		}
	}

	if (use_graph) {

		if (graph_iters != maxiter || graph_lambda != lambda) {

			drop_graph();
			BAM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
			prof_set_capturing(true);

			try {
				for (int i = 0; i < maxiter; i++)
					enqueue_iteration(normal, lambda);
			} catch (...) {
				prof_set_capturing(false);
				(void)hipStreamEndCapture(stream, &graph);
				throw;
			}

			prof_set_capturing(false);
			BAM_HIP(hipStreamEndCapture(stream, &graph));
			BAM_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
			graph_iters = maxiter;
			graph_lambda = lambda;
		}

		BAM_HIP(hipGraphLaunch(gexec, stream));

	} else {

		for (int i = 0; i < maxiter; i++)
			enqueue_iteration(normal, lambda);
	}

	BAM_HIP(hipMemcpyAsync(x_io, x.ptr, sizeof(float) * n, hipMemcpyDeviceToDevice, stream));
	BAM_HIP(hipMemcpyAsync(host_st, st.ptr, sizeof(CgState), hipMemcpyDeviceToHost, stream));
	BAM_HIP(hipEventRecord(e1, stream));
	BAM_HIP(hipStreamSynchronize(stream));

	float ms = 0.f;
	BAM_HIP(hipEventElapsedTime(&ms, e0, e1));
	BAM_HIP(hipEventDestroy(e0));
	BAM_HIP(hipEventDestroy(e1));

	res.iterations = host_st->iter;
	res.resid = sqrtf(host_st->rsnew);
	res.bnorm = host_st->bnorm;
	res.status = host_st->done;
	res.ms = ms;

	if (res.iterations > 0) {
		res.trace.resize(res.iterations);
		BAM_HIP(hipMemcpy(res.trace.data(), trace.ptr, sizeof(float) * res.iterations, hipMemcpyDeviceToHost));
	}

	return res;
}

}  // namespace bam
