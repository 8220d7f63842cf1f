// cg.hpp -- device-resident conjugate-gradient solver.
#pragma once

#include "nufft_impl.hpp"

#include <functional>
#include <vector>

namespace bam {

struct CgState {

	float rsold;
	float rsnew;
	float alpha;
	float beta;
	float eps2;
	float bnorm;
	float r0norm;
	int iter;
	int maxiter;
	int done;  // 0 running, 1 maxiter, 2 converged, 3 pAp == 0, 4 nothing to do
};

struct CgResult {

	int iterations = 0;
	float resid = 0.f;   // ||r|| (recursively updated residual)
	float bnorm = 0.f;   // ||b||
	int status = 0;
	bool fell_back = false;
	float ms = 0.f;
	std::vector<float> trace;
};

// normal(dst, src, skip): enqueue dst = A^H A src on the solver's stream;
// skip is a device flag (may be null) that turns the launches into no-ops.
using NormalFn = std::function<void(float2*, const float2*, const int*)>;

struct CgSolver {

	long n;  // floats (2 x complex elements)
	hipStream_t stream;
	DevBuf<float> x, b, r, p, Ap, trace;
	DevBuf<double> part, part_b;
	DevBuf<CgState> st;
	CgState* host_st = nullptr;
	int nblocks = 1;

	hipGraph_t graph = nullptr;
	hipGraphExec_t gexec = nullptr;
	int graph_iters = -1;
	float graph_lambda = 0.f;

	CgSolver(long nfloats, hipStream_t s);
	~CgSolver();
	CgSolver(const CgSolver&) = delete;
	CgSolver& operator=(const CgSolver&) = delete;

	void drop_graph();
	void enqueue_iteration(const NormalFn& normal, float lambda);

	// x_io, b: device pointers of n floats.  warm: x_io holds the initial
	// guess; fallback: restart from zero when the warm start's residual
	// exceeds ||b||.
	CgResult solve(const NormalFn& normal, float lambda, int maxiter, float tol, float2* x_io, const float2* b,
		       bool warm, bool fallback, bool use_graph);
};

}  // namespace bam
