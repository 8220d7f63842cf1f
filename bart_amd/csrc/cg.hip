// cg.hip -- conjugate gradients on (A^H A + lambda I) x = b, gfx950, for one
// problem or a batch of independent problems.
//
// Arithmetic of iter/italgos.c:497-580 (conjgrad) with the stopping rule of
// iter/iter2.c:227-270 (stop when ||r||^2 <= (tol ||b||)^2; tol = 0 runs
// maxiter iterations) and the FP64 dot/norm accumulation of
// num/vecops.c:89-124.  A batch packs the vectors of its problems into one
// buffer each and keeps the scalars per problem, as conjgrad_batch
// (iter/italgos.c:589-680, dot_bat/axpy_bat/xpay_bat of
// num/gpukrnls_bat.cu) does; every problem stops by its own test, so each
// one follows exactly the iterations of its single solve.  Differences in
// execution, not in maths:
//   * the scalars (rsold, alpha, beta, stop flags) stay on the device; the
//     host does not read anything until the solve ends (BART's GPU path copies
//     two scalars to the host every iteration, num/gpukrnls.cu:1847-1858);
//   * reductions are two-stage and deterministic (per-work-group FP64
//     partials, then a fixed-order sum per problem);
//   * one iteration -- normal operators and vector updates -- is captured
//     once into a hipGraph and launched up to maxiter times, in chunks
//     whose stop flags the host reads one chunk behind; after a problem
//     stops, its kernels see the device stop flag and return immediately.
#include "cg.hpp"
#include "prof.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace bam {

constexpr int CG_THREADS = 256;
constexpr int CG_BLOCKS = 512;   // work-groups per problem at most
constexpr int CG_CHUNK = 8;      // iterations launched between stop-flag reads

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

// the element range of work-group b of its problem: first element and stride
struct SegIt {

	long i0, step, len, off;
};

__device__ inline SegIt seg_it(const CgSeg& s, int b)
{
	SegIt it;
	it.i0 = (long)(b - s.blk0) * blockDim.x + threadIdx.x;
	it.step = (long)s.nblk * blockDim.x;
	it.len = s.len;
	it.off = s.off;
	return it;
}

// the sum of a problem's partials in fixed order, in every thread of the
// work-group (each work-group of a vector launch reduces them itself, so
// that no separate single-work-group launch is needed and no atomics)
__device__ inline double sum_partials_all(const double* part, int n, double* sh, double* bc)
{
	double v = sum_partials(part, n, sh);

	if (0 == threadIdx.x)
		*bc = v;

	__syncthreads();
	return *bc;
}

// Ap += lambda p; partial sums of p . Ap (float products, FP64 accumulation).
// Work-group b = blk_base + blockIdx.x.
__global__ __launch_bounds__(CG_THREADS) void k_cg_apdot(const int* __restrict__ bprob, const CgSeg* __restrict__ seg,
							 float lambda, const float* __restrict__ p,
							 float* __restrict__ Ap, double* __restrict__ part,
							 const CgState* st, int blk_base)
{
	__shared__ double sh[4];
	const int b = blk_base + (int)blockIdx.x;
	const int q = bprob[b];

	if (0 != st[q].done)
		return;

	const SegIt it = seg_it(seg[q], b);
	p += it.off;
	Ap += it.off;
	double acc = 0.;

	for (long i = it.i0; i < it.len; i += it.step) {

		float pv = p[i];
		float a = Ap[i] + lambda * pv;
		Ap[i] = a;
		acc += (double)(pv * a);
	}

	double s = block_sum(acc, sh);

	if (0 == threadIdx.x)
		part[b] = s;
}

// The same step fused into the sum of the coil-group partial images that
// ends a normal operator: Ap = sum_g grp[g] + lambda p (groups added in
// order, as k_sum_groups), partial sums of p . Ap in the layout of
// k_cg_apdot.
__global__ __launch_bounds__(CG_THREADS) void k_cg_apdot_groups(CgApFuse f, const float* __restrict__ grp,
								long gstride, int G)
{
	__shared__ double sh[4];

	if (nullptr != f.skip && 0 != *f.skip)
		return;

	const long step = (long)f.nblk * blockDim.x;
	double acc = 0.;

	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < f.len; i += step) {

		float a = grp[i];

		for (int g = 1; g < G; g++)
			a = a + grp[(long)g * gstride + i];

		float pv = f.p[i];
		float v = a + f.lambda * pv;
		f.Ap[i] = v;
		acc += (double)(pv * v);
	}

	double s = block_sum(acc, sh);

	if (0 == threadIdx.x)
		f.part[f.blk0 + blockIdx.x] = s;
}

// The same step for every problem of a batch whose normal operator ran
// grouped (items[q].grp != null), in the layout of k_cg_apdot.
__global__ __launch_bounds__(CG_THREADS) void k_cg_apdot_sum_batch(const int* __restrict__ bprob,
								   const CgSeg* __restrict__ seg,
								   const CgGrpItem* __restrict__ items, float lambda,
								   const float* __restrict__ p, float* __restrict__ Ap,
								   double* __restrict__ part, const CgState* st)
{
	__shared__ double sh[4];
	const int b = (int)blockIdx.x;
	const int q = bprob[b];
	const CgGrpItem it = items[q];

	if (nullptr == it.grp || 0 != st[q].done)
		return;

	const SegIt si = seg_it(seg[q], b);
	p += si.off;
	Ap += si.off;
	double acc = 0.;

	for (long i = si.i0; i < si.len; i += si.step) {

		float a = it.grp[i];

		for (int g = 1; g < it.G; g++)
			a = a + it.grp[(long)g * it.gstride + i];

		float pv = p[i];
		float v = a + lambda * pv;
		Ap[i] = v;
		acc += (double)(pv * v);
	}

	double s = block_sum(acc, sh);

	if (0 == threadIdx.x)
		part[b] = s;
}

void cg_apdot_groups(const CgApFuse& f, const float2* grp, long N, int G, hipStream_t st)
{
	ProfScope prof("cg_apdot_groups", st);
	hipLaunchKernelGGL(k_cg_apdot_groups, dim3(f.nblk), dim3(CG_THREADS), 0, st, f, (const float*)grp, 2 * N, G);
	BAM_HIP(hipGetLastError());
}

// alpha = rsnew / p.Ap (every work-group reduces the p.Ap partials itself;
// the first one records alpha, copies rsnew to rsold for the beta step and
// flags a breakdown, pAp = 0, italgos.c:544-545); x += alpha p;
// r -= alpha Ap; partial sums of r . r
__global__ __launch_bounds__(CG_THREADS) void k_cg_xr(const int* __restrict__ bprob, const CgSeg* __restrict__ seg,
						      float* __restrict__ x, float* __restrict__ r,
						      const float* __restrict__ p, const float* __restrict__ Ap,
						      const double* __restrict__ part_ap, double* __restrict__ part_rr,
						      CgState* st, CgMon mon)
{
	__shared__ double sh[4];
	__shared__ double bc;
	__shared__ float s_alpha;
	const int q = bprob[blockIdx.x];

	if (0 != st[q].done)
		return;

	const CgSeg sg = seg[q];
	const double s = sum_partials_all(part_ap + sg.blk0, sg.nblk, sh, &bc);
	const float pAp = (float)s;

	if (0.f == pAp) {
		if (0 == threadIdx.x && (int)blockIdx.x == sg.blk0)
			st[q].done = 3;     // x unchanged
		return;
	}

	if (0 == threadIdx.x) {
		const float rs = st[q].rsnew;
		s_alpha = rs / pAp;
		if ((int)blockIdx.x == sg.blk0) {
			st[q].alpha = s_alpha;
			st[q].rsold = rs;
		}
	}

	__syncthreads();
	const float alpha = s_alpha;
	const SegIt it = seg_it(sg, (int)blockIdx.x);
	x += it.off;
	r += it.off;
	p += it.off;
	Ap += it.off;
	double acc = 0.;

	if (nullptr != mon.truth && 0.f != mon.tinv[q]) {

		// the monitor sees the iterate this iteration starts from
		const float* t = mon.truth + it.off;
		double acce = 0.;

		for (long i = it.i0; i < it.len; i += it.step) {

			const float xo = x[i];
			const float d = t[i] - xo;
			acce += (double)(d * d);
			x[i] = xo + alpha * p[i];
			float rv = r[i] + (-alpha) * Ap[i];
			r[i] = rv;
			acc += (double)(rv * rv);
		}

		__syncthreads();
		double se = block_sum(acce, sh);

		if (0 == threadIdx.x)
			mon.part[blockIdx.x] = se;

	} else {

		for (long i = it.i0; i < it.len; i += it.step) {

			x[i] += alpha * p[i];
			float rv = r[i] + (-alpha) * Ap[i];
			r[i] = rv;
			acc += (double)(rv * rv);
		}
	}

	__syncthreads();
	double s2 = block_sum(acc, sh);

	if (0 == threadIdx.x)
		part_rr[blockIdx.x] = s2;
}

// beta = rsnew / rsold (every work-group reduces the r.r partials itself;
// the first one stores rsnew, the iteration count, the trace entry
// ([B][tstride] residual norms) and the stop flag); p = beta p + r
__global__ __launch_bounds__(CG_THREADS) void k_cg_p(const int* __restrict__ bprob, const CgSeg* __restrict__ seg,
						     float* __restrict__ p, const float* __restrict__ r,
						     const double* __restrict__ part_rr, CgState* st,
						     float* __restrict__ trace, int tstride, CgMon mon)
{
	__shared__ double sh[4];
	__shared__ double bc;
	__shared__ float s_beta;
	const int q = bprob[blockIdx.x];

	if (0 != st[q].done)
		return;

	const CgSeg sg = seg[q];

	if ((int)blockIdx.x == sg.blk0 && nullptr != mon.truth && 0.f != mon.tinv[q]) {

		const double se = sum_partials(mon.part + sg.blk0, sg.nblk, sh);

		if (0 == threadIdx.x && st[q].iter < tstride)
			mon.err[(long)q * tstride + st[q].iter] = (float)(sqrt(se) * (double)mon.tinv[q]);

		__syncthreads();
	}

	const double s = sum_partials_all(part_rr + sg.blk0, sg.nblk, sh, &bc);

	if (0 == threadIdx.x) {

		const double nr = sqrt(s);
		const float rsnew = (float)(nr * nr);
		s_beta = rsnew / st[q].rsold;

		if ((int)blockIdx.x == sg.blk0) {

			CgState* sq = st + q;
			const int itn = sq->iter;

			sq->beta = s_beta;
			sq->rsnew = rsnew;
			sq->iter = itn + 1;

			if (itn < tstride)
				trace[(long)q * tstride + itn] = sqrtf(rsnew);

			if (rsnew <= sq->eps2)
				sq->done = 2;  // converged
			else if (sq->iter >= sq->maxiter)
				sq->done = 1;  // iteration budget used
		}
	}

	__syncthreads();
	const float beta = s_beta;
	const SegIt it = seg_it(sg, (int)blockIdx.x);
	p += it.off;
	r += it.off;

	for (long i = it.i0; i < it.len; i += it.step)
		p[i] = beta * p[i] + r[i];
}

// r = b - (A x + lambda x) [with ax = A x already in r]; p = r; ||r||^2 and
// ||b||^2 partials.  mask: only the problems with mask[q] != 0 (null: all).
__global__ __launch_bounds__(CG_THREADS) void k_cg_init(const int* __restrict__ bprob, const CgSeg* __restrict__ seg,
							float lambda, const float* __restrict__ x,
							const float* __restrict__ b, float* __restrict__ r,
							float* __restrict__ p, double* __restrict__ part,
							double* __restrict__ part_b, bool cold, const int* mask)
{
	__shared__ double sh[4];
	const int q = bprob[blockIdx.x];

	if (nullptr != mask && 0 == mask[q])
		return;

	const SegIt it = seg_it(seg[q], (int)blockIdx.x);
	x += it.off;
	b += it.off;
	r += it.off;
	p += it.off;
	double acc = 0., accb = 0.;

	for (long i = it.i0; i < it.len; i += it.step) {

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

// one work-group per problem
__global__ void k_cg_init_fin(const CgSeg* __restrict__ seg, const double* part, const double* part_b, float tol,
			      int maxiter, CgState* st, const int* mask)
{
	__shared__ double sh[4];
	const int q = blockIdx.x;

	if (nullptr != mask && 0 == mask[q])
		return;

	double s = sum_partials(part + seg[q].blk0, seg[q].nblk, sh);
	__syncthreads();
	double sb = sum_partials(part_b + seg[q].blk0, seg[q].nblk, sh);

	if (0 == threadIdx.x) {

		CgState* sq = st + q;
		double nb = sqrt(sb);
		float eps = (float)nb * tol;
		double nr = sqrt(s);
		float rs = (float)(nr * nr);

		sq->bnorm = (float)nb;
		sq->r0norm = (float)nr;
		sq->rsold = rs;
		sq->rsnew = rs;
		sq->eps2 = (float)pow((double)eps, 2.);
		sq->iter = 0;
		sq->maxiter = maxiter;
		sq->alpha = 0.f;
		sq->beta = 0.f;
		// checkeps (iter2.c:65-80) and the early out of italgos.c:526-530
		sq->done = (0. == nb || 0.f == rs || maxiter <= 0) ? 4 : 0;
	}
}

// ---------------------------------------------------------------------------

CgBatch::CgBatch(const std::vector<long>& nfloats, hipStream_t s) : B((int)nfloats.size()), stream(s)
{
	BAM_CHECK(B >= 1, "cg: empty batch");

	seg_h.resize(B);
	std::vector<int> bp;

	for (int i = 0; i < B; i++) {

		long n = nfloats[i];
		BAM_CHECK(n >= 0, "cg: negative problem size");
		CgSeg& sg = seg_h[i];
		sg.off = total;
		sg.len = n;
		sg.blk0 = nblk;
		sg.nblk = (int)std::min<long>(CG_BLOCKS, std::max<long>(1, (n + CG_THREADS - 1) / CG_THREADS));
		nblk += sg.nblk;
		total += (n + 63) / 64 * 64;  // 256-byte aligned problems
		bp.insert(bp.end(), sg.nblk, i);
	}

	total = std::max<long>(total, 64);
	x.alloc(total);
	b.alloc(total);
	r.alloc(total);
	p.alloc(total);
	Ap.alloc(total);
	part.alloc(nblk);
	part_b.alloc(nblk);
	part2.alloc(nblk);
	seg.alloc(B);
	bprob.alloc(nblk);
	mask.alloc(B);
	st.alloc(B);
	trace.alloc(B);
	trace_len = 1;
	fell_back.assign(B, 0);

	BAM_HIP(hipMemcpy(seg.ptr, seg_h.data(), sizeof(CgSeg) * B, hipMemcpyHostToDevice));
	BAM_HIP(hipMemcpy(bprob.ptr, bp.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
	BAM_HIP(hipHostMalloc((void**)&host_st, sizeof(CgState) * B, hipHostMallocDefault));
	BAM_HIP(hipHostMalloc((void**)&probe, 2 * sizeof(CgState) * B, hipHostMallocDefault));
	BAM_HIP(hipEventCreateWithFlags(&probe_ev[0], hipEventDisableTiming));
	BAM_HIP(hipEventCreateWithFlags(&probe_ev[1], hipEventDisableTiming));
}

CgBatch::~CgBatch()
{
	drop_graph();
	if (nullptr != host_st)
		(void)hipHostFree(host_st);
	if (nullptr != probe)
		(void)hipHostFree(probe);
	for (hipEvent_t e : probe_ev)
		if (nullptr != e)
			(void)hipEventDestroy(e);
}

CgMon CgBatch::mon() const
{
	if (!monitored)
		return CgMon{};
	return CgMon{ mon_truth.ptr, mon_tinv.ptr, mon_part.ptr, mon_err.ptr };
}

void CgBatch::set_monitor(const float* truth, double tnorm)
{
	BAM_CHECK(1 == B, "cg monitor: single problems only");
	drop_graph();  // the iteration's kernels take the monitor's pointers

	if (nullptr == truth) {
		monitored = false;
		return;
	}

	const long n = seg_h[0].len;
	if ((long)mon_truth.n < n)
		mon_truth.alloc(n);
	if ((long)mon_part.n < nblk)
		mon_part.alloc(nblk);
	if (mon_tinv.n < 1)
		mon_tinv.alloc(1);
	BAM_HIP(hipMemcpy(mon_truth.ptr, truth, sizeof(float) * n, hipMemcpyDeviceToDevice));
	const float tinv = (tnorm > 0.) ? (float)(1. / tnorm) : 0.f;
	BAM_HIP(hipMemcpy(mon_tinv.ptr, &tinv, sizeof(float), hipMemcpyHostToDevice));
	monitored = true;
}

void CgBatch::monitor_errors(float* host, int n) const
{
	if (monitored && n > 0)
		BAM_HIP(hipMemcpy(host, mon_err.ptr, sizeof(float) * n, hipMemcpyDeviceToHost));
}

void CgBatch::drop_graph()
{
	if (nullptr != gexec)
		(void)hipGraphExecDestroy(gexec);
	if (nullptr != graph)
		(void)hipGraphDestroy(graph);
	gexec = nullptr;
	graph = nullptr;
	have_graph = false;
}

void CgBatch::init_kernels(float lambda, float tol, int maxiter, bool cold, bool masked)
{
	const int* m = masked ? mask.ptr : nullptr;
	hipLaunchKernelGGL(k_cg_init, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob.ptr, seg.ptr, lambda, x.ptr, b.ptr,
			   r.ptr, p.ptr, part.ptr, part_b.ptr, cold, m);
	hipLaunchKernelGGL(k_cg_init_fin, dim3(B), dim3(CG_THREADS), 0, stream, seg.ptr, part.ptr, part_b.ptr, tol,
			   maxiter, st.ptr, m);
	BAM_HIP(hipGetLastError());
}

void CgBatch::start(const BatchNormalFn& normal, float lambda, float tol, int maxiter, const float2* const* x0,
		    const float2* const* b_in, bool warm, bool fallback)
{
	std::fill(fell_back.begin(), fell_back.end(), 0);

	// per-problem residual trace, allocated before any capture
	if (trace_len < std::max(1, maxiter)) {
		trace_len = std::max(1, maxiter);
		trace.alloc((size_t)B * trace_len);
		drop_graph();
	}

	if (monitored && (long)mon_err.n < (long)B * trace_len) {
		mon_err.alloc((size_t)B * trace_len);
		drop_graph();
	}

	for (int i = 0; i < B; i++)
		BAM_HIP(hipMemcpyAsync(bs(i), b_in[i], sizeof(float) * seg_h[i].len, hipMemcpyDeviceToDevice, stream));

	if (warm) {

		for (int i = 0; i < B; i++) {
			BAM_HIP(hipMemcpyAsync(xs(i), x0[i], sizeof(float) * seg_h[i].len, hipMemcpyDeviceToDevice, stream));
			normal(i, (float2*)rs(i), (const float2*)xs(i), nullptr, nullptr);
		}

	} else {

		fill((float2*)x.ptr, total / 2, make_float2(0.f, 0.f), stream);
	}

	init_kernels(lambda, tol, maxiter, !warm, false);

	if (warm && fallback) {

		// zero start for every problem whose warm-start residual exceeds ||b||
		BAM_HIP(hipMemcpyAsync(host_st, st.ptr, sizeof(CgState) * B, hipMemcpyDeviceToHost, stream));
		BAM_HIP(hipStreamSynchronize(stream));

		std::vector<int> m(B, 0);
		int any = 0;

		for (int i = 0; i < B; i++) {
			if (host_st[i].r0norm > host_st[i].bnorm) {
				m[i] = 1;
				fell_back[i] = 1;
				any = 1;
			}
		}

		if (any) {

			BAM_HIP(hipMemcpyAsync(mask.ptr, m.data(), sizeof(int) * B, hipMemcpyHostToDevice, stream));

			for (int i = 0; i < B; i++)
				if (m[i])
					fill((float2*)xs(i), seg_h[i].len / 2, make_float2(0.f, 0.f), stream);

			init_kernels(lambda, tol, maxiter, true, true);
			BAM_HIP(hipStreamSynchronize(stream));  // m leaves scope
		}
	}
}

void CgBatch::enqueue_iteration(const BatchNormalFn& normal, float lambda)
{
	// A p, and where the operator ends with a coil-group sum, the step
	// A p += lambda p with its p . Ap partials fused into that sum
	std::vector<char> fused(B, 0);
	int nfused = 0;

	// the problems in groups: their passes, then one launch of the step
	if (nullptr != hooks && hooks_ready) {
		hooks->launch(*this);
		hipLaunchKernelGGL(k_cg_apdot_sum_batch, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob.ptr, seg.ptr,
				   grp.ptr, lambda, p.ptr, Ap.ptr, part.ptr, st.ptr);
		for (int i = 0; i < B; i++)
			if (nullptr != grp_h[i].grp) {
				fused[i] = 1;
				nfused++;
			}
	}

	for (int i = 0; i < B; i++) {
		if (fused[i])
			continue;
		const CgSeg& sg = seg_h[i];
		CgApFuse f{ lambda, p.ptr + sg.off, Ap.ptr + sg.off, part.ptr, sg.blk0, sg.nblk, sg.len, &st.ptr[i].done };
		fused[i] = normal(i, (float2*)(Ap.ptr + sg.off), (const float2*)(p.ptr + sg.off), &st.ptr[i].done, &f);
		nfused += fused[i] ? 1 : 0;
	}

	ProfScope prof("cg_vector_ops", stream);

	if (0 == nfused) {
		hipLaunchKernelGGL(k_cg_apdot, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob.ptr, seg.ptr, lambda, p.ptr,
				   Ap.ptr, part.ptr, st.ptr, 0);
	} else {
		for (int i = 0; i < B; i++)
			if (!fused[i])
				hipLaunchKernelGGL(k_cg_apdot, dim3(seg_h[i].nblk), dim3(CG_THREADS), 0, stream, bprob.ptr,
						   seg.ptr, lambda, p.ptr, Ap.ptr, part.ptr, st.ptr, seg_h[i].blk0);
	}

	const CgMon m = mon();
	hipLaunchKernelGGL(k_cg_xr, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob.ptr, seg.ptr, x.ptr, r.ptr, p.ptr,
			   Ap.ptr, part.ptr, part2.ptr, st.ptr, m);
	hipLaunchKernelGGL(k_cg_p, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob.ptr, seg.ptr, p.ptr, r.ptr, part2.ptr,
			   st.ptr, trace.ptr, trace_len, m);
	BAM_HIP(hipGetLastError());
}

void CgBatch::capture(const BatchNormalFn& normal, float lambda)
{
	if (have_graph && graph_lambda == lambda)
		return;

	drop_graph();

	// tables of the grouped normal operators (device memory: before the capture)
	if (nullptr != hooks && !hooks_ready) {
		grp_h.assign(B, CgGrpItem{ nullptr, 0, 0 });
		hooks->prepare(*this, grp_h);
		grp.alloc(B);
		BAM_HIP(hipMemcpy(grp.ptr, grp_h.data(), sizeof(CgGrpItem) * B, hipMemcpyHostToDevice));
		hooks_ready = true;
	}
	BAM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
	prof_set_capturing(true);

	try {
		enqueue_iteration(normal, lambda);
	} catch (...) {
		prof_set_capturing(false);
		(void)hipStreamEndCapture(stream, &graph);
		throw;
	}

	prof_set_capturing(false);
	BAM_HIP(hipStreamEndCapture(stream, &graph));
	BAM_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
	have_graph = true;
	graph_lambda = lambda;
}

void CgBatch::launch(int n)
{
	for (int i = 0; i < n; i++)
		BAM_HIP(hipGraphLaunch(gexec, stream));
}

void CgBatch::post_probe(int slot)
{
	BAM_HIP(hipMemcpyAsync(probe + (size_t)slot * B, st.ptr, sizeof(CgState) * B, hipMemcpyDeviceToHost, stream));
	BAM_HIP(hipEventRecord(probe_ev[slot], stream));
}

bool CgBatch::probe_all_done(int slot)
{
	BAM_HIP(hipEventSynchronize(probe_ev[slot]));
	const CgState* ps = probe + (size_t)slot * B;

	for (int q = 0; q < B; q++)
		if (0 == ps[q].done)
			return false;

	return true;
}

// Launch in chunks, round-robin over the batches (one stream each); after a
// batch's chunk its stop flags are copied to the host, and the copy of the
// chunk before is read while the current one runs, so no device waits for
// the host.  A batch whose problems have all stopped gets no more launches
// (at most one chunk of no-op iterations follows its last useful one).
void iterate_batches(const std::vector<CgBatch*>& bs, const std::vector<const BatchNormalFn*>& fns, float lambda,
		     int maxiter)
{
	const int K = (int)bs.size();

	if (maxiter <= 0 || 0 == K)
		return;

	for (int k = 0; k < K; k++)
		bs[k]->capture(*fns[k], lambda);

	std::vector<int> launched(K, 0);
	std::vector<char> active(K, 1);
	int left = K;

	for (int chunk = 0; left > 0; chunk++) {

		const int slot = chunk & 1;

		for (int k = 0; k < K; k++) {

			if (!active[k])
				continue;

			const int nl = std::min(CG_CHUNK, maxiter - launched[k]);
			bs[k]->launch(nl);
			launched[k] += nl;

			if (launched[k] >= maxiter) {
				active[k] = 0;
				left--;
			} else {
				bs[k]->post_probe(slot);
			}
		}

		if (chunk > 0)
			for (int k = 0; k < K; k++)
				if (active[k] && bs[k]->probe_all_done(1 - slot)) {
					active[k] = 0;
					left--;
				}
	}
}

void CgBatch::iterate(const BatchNormalFn& normal, float lambda, int maxiter, bool use_graph)
{
	if (maxiter <= 0)
		return;

	if (!use_graph) {

		for (int i = 0; i < maxiter; i++)
			enqueue_iteration(normal, lambda);
		return;
	}

	iterate_batches({ this }, { &normal }, lambda, maxiter);
}

void CgBatch::finish(float2* const* x_out)
{
	for (int i = 0; i < B; i++)
		BAM_HIP(hipMemcpyAsync(x_out[i], xs(i), sizeof(float) * seg_h[i].len, hipMemcpyDeviceToDevice, stream));

	BAM_HIP(hipMemcpyAsync(host_st, st.ptr, sizeof(CgState) * B, hipMemcpyDeviceToHost, stream));
}

CgResult CgBatch::result(int i) const
{
	CgResult res;
	const CgState& s = host_st[i];

	res.iterations = s.iter;
	res.resid = sqrtf(s.rsnew);
	res.bnorm = s.bnorm;
	res.status = s.done;
	res.fell_back = 0 != fell_back[i];

	if (res.iterations > 0) {
		res.trace.resize(res.iterations);
		BAM_HIP(hipMemcpy(res.trace.data(), trace.ptr + (long)i * trace_len, sizeof(float) * res.iterations,
				  hipMemcpyDeviceToHost));
	}

	return res;
}

// ---------------------------------------------------------------------------
// Allocation-free single solve
// ---------------------------------------------------------------------------

// Working memory of cg_solve_noalloc -- r, p and A p of one problem of at most
// CG_NOALLOC_ELEMS complex elements, the reduction partials, the state, the
// segment tables and the residual trace -- in one array of this library.
// Every GPU the library runs on holds its own copy (48 MiB + 64 KiB) until the
// process ends; every allocation-free solve on that GPU shares it, so such
// calls that overlap on one GPU race.
constexpr size_t NA_VEC = (size_t)CG_NOALLOC_ELEMS * sizeof(float2);
constexpr size_t NA_PART = NA_VEC * 3;                                 // double[CG_BLOCKS]
constexpr size_t NA_PART_B = NA_PART + CG_BLOCKS * sizeof(double);     // double[CG_BLOCKS]
constexpr size_t NA_ST = NA_PART_B + CG_BLOCKS * sizeof(double);       // CgState
constexpr size_t NA_SEG = NA_ST + 256;                                 // CgSeg
constexpr size_t NA_BPROB = NA_SEG + 256;                              // int[CG_BLOCKS]
constexpr size_t NA_PART2 = NA_BPROB + CG_BLOCKS * sizeof(int);        // double[CG_BLOCKS]
constexpr size_t NA_TRACE = NA_PART2 + CG_BLOCKS * sizeof(double);     // float[CG_NOALLOC_TRACE]
constexpr size_t NA_BYTES = NA_VEC * 3 + 65536;
static_assert(NA_TRACE + CG_NOALLOC_TRACE * sizeof(float) <= NA_BYTES, "allocation-free CG layout");

static __device__ __align__(256) char scratch[NA_BYTES];

// device address of this GPU's copy of scratch (the host-side name is a
// stand-in, not a device address), looked up once per device
static std::mutex scratch_mtx;
static void* scratch_dev[64] = {};

static char* scratch_base()
{
	int dev = 0;
	BAM_HIP(hipGetDevice(&dev));
	BAM_CHECK(dev >= 0 && dev < 64, "cg: device index %d out of range", dev);
	std::lock_guard<std::mutex> lock(scratch_mtx);

	if (nullptr == scratch_dev[dev]) {

		void* p = nullptr;
		BAM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		BAM_CHECK(nullptr != p, "cg: no device address for the solver's working memory");
		scratch_dev[dev] = p;
	}

	return (char*)scratch_dev[dev];
}

// the segment table of one problem of len floats over nblk work-groups
__global__ void k_cg_single_tables(CgSeg* seg, int* bprob, long len, int nblk)
{
	if (0 == threadIdx.x && 0 == blockIdx.x) {
		seg->off = 0;
		seg->len = len;
		seg->blk0 = 0;
		seg->nblk = nblk;
	}

	for (int i = threadIdx.x; i < nblk; i += blockDim.x)
		bprob[i] = 0;
}

CgResult cg_solve_noalloc(const NormalFn& normal, float lambda, int maxiter, float tol, float2* x,
			  const float2* b, long nfloats, bool warm, bool fallback, hipStream_t stream,
			  float* trace_host)
{
	BAM_CHECK(nfloats >= 0 && nfloats <= 2 * CG_NOALLOC_ELEMS,
		  "lsqr_conjgrad_noalloc: image of %ld elements exceeds the maximum of %ld", nfloats / 2,
		  CG_NOALLOC_ELEMS);
	BAM_CHECK(maxiter <= CG_NOALLOC_TRACE, "lsqr_conjgrad_noalloc: %d iterations exceed the maximum of %d",
		  maxiter, CG_NOALLOC_TRACE);

	char* base = scratch_base();
	float* r = (float*)base;
	float* p = (float*)(base + NA_VEC);
	float* Ap = (float*)(base + 2 * NA_VEC);
	double* part = (double*)(base + NA_PART);
	double* part_b = (double*)(base + NA_PART_B);
	CgState* st = (CgState*)(base + NA_ST);
	CgSeg* seg = (CgSeg*)(base + NA_SEG);
	int* bprob = (int*)(base + NA_BPROB);
	double* part2 = (double*)(base + NA_PART2);
	float* trace = (float*)(base + NA_TRACE);
	float* xf = (float*)x;
	const float* bf = (const float*)b;

	// the blocking of a batch of one (CgBatch), so that the results equal
	// those of CgSolver bit for bit
	const int nblk = (int)std::min<long>(CG_BLOCKS, std::max<long>(1, (nfloats + CG_THREADS - 1) / CG_THREADS));

	hipLaunchKernelGGL(k_cg_single_tables, dim3(1), dim3(CG_THREADS), 0, stream, seg, bprob, nfloats, nblk);

	auto init = [&](bool cold) {
		hipLaunchKernelGGL(k_cg_init, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob, seg, lambda, xf, bf, r, p,
				   part, part_b, cold, (const int*)nullptr);
		hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(CG_THREADS), 0, stream, seg, part, part_b, tol, maxiter,
				   st, (const int*)nullptr);
		BAM_HIP(hipGetLastError());
	};

	CgResult res;
	CgState hs;

	if (warm) {
		normal((float2*)r, x, nullptr, nullptr);
		init(false);
	} else {
		fill(x, nfloats / 2, make_float2(0.f, 0.f), stream);
		init(true);
	}

	if (warm && fallback) {

		BAM_HIP(hipMemcpyAsync(&hs, st, sizeof(CgState), hipMemcpyDeviceToHost, stream));
		BAM_HIP(hipStreamSynchronize(stream));

		if (hs.r0norm > hs.bnorm) {
			fill(x, nfloats / 2, make_float2(0.f, 0.f), stream);
			init(true);
			res.fell_back = true;
		}
	}

	// iterations in chunks; the stop flag is read after each chunk (the
	// kernels of a stopped problem return at once)
	for (int done = 0; done < maxiter;) {

		const int nl = std::min(CG_CHUNK, maxiter - done);

		for (int i = 0; i < nl; i++) {
			CgApFuse f{ lambda, p, Ap, part, 0, nblk, nfloats, &st->done };
			if (!normal((float2*)Ap, (const float2*)p, &st->done, &f))
				hipLaunchKernelGGL(k_cg_apdot, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob, seg, lambda, p,
						   Ap, part, (const CgState*)st, 0);
			hipLaunchKernelGGL(k_cg_xr, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob, seg, xf, r, p, Ap, part,
					   part2, st, CgMon{});
			hipLaunchKernelGGL(k_cg_p, dim3(nblk), dim3(CG_THREADS), 0, stream, bprob, seg, p, r, part2, st, trace,
					   CG_NOALLOC_TRACE, CgMon{});
			BAM_HIP(hipGetLastError());
		}

		done += nl;

		BAM_HIP(hipMemcpyAsync(&hs, st, sizeof(CgState), hipMemcpyDeviceToHost, stream));
		BAM_HIP(hipStreamSynchronize(stream));

		if (0 != hs.done)
			break;
	}

	BAM_HIP(hipMemcpyAsync(&hs, st, sizeof(CgState), hipMemcpyDeviceToHost, stream));
	BAM_HIP(hipStreamSynchronize(stream));

	res.iterations = hs.iter;
	res.resid = sqrtf(hs.rsnew);
	res.bnorm = hs.bnorm;
	res.status = hs.done;

	if (nullptr != trace_host && res.iterations > 0)
		BAM_HIP(hipMemcpy(trace_host, trace, sizeof(float) * res.iterations, hipMemcpyDeviceToHost));

	return res;
}

CgResult CgSolver::solve(const NormalFn& normal, float lambda, int maxiter, float tol, float2* x_io,
			 const float2* b_in, bool warm, bool fallback, bool use_graph)
{
	BatchNormalFn bn = [&normal](int, float2* d, const float2* s, const int* skip, const CgApFuse* f) {
		return normal(d, s, skip, f);
	};

	hipEvent_t e0, e1;
	BAM_HIP(hipEventCreate(&e0));
	BAM_HIP(hipEventCreate(&e1));
	BAM_HIP(hipEventRecord(e0, bt.stream));

	const float2* x0 = x_io;
	bt.start(bn, lambda, tol, maxiter, &x0, &b_in, warm, fallback);
	bt.iterate(bn, lambda, maxiter, use_graph);
	bt.finish(&x_io);

	BAM_HIP(hipEventRecord(e1, bt.stream));
	BAM_HIP(hipStreamSynchronize(bt.stream));

	float ms = 0.f;
	BAM_HIP(hipEventElapsedTime(&ms, e0, e1));
	BAM_HIP(hipEventDestroy(e0));
	BAM_HIP(hipEventDestroy(e1));

	CgResult res = bt.result(0);
	res.ms = ms;
	return res;
}

}  // namespace bam
