// cg.hpp -- device-resident conjugate-gradient solver for one or many
// independent problems.
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

// One problem's part of the packed vectors: floats [off, off + len), reduced
// by the work-groups [blk0, blk0 + nblk) of a vector launch.
struct CgSeg {

	long off;
	long len;
	int blk0;
	int nblk;
};

// The CG step that follows A^H A p -- Ap = A^H A p + lambda p and the
// partial sums of p . Ap -- for a normal operator whose last kernel can take
// it (the sum of coil-group partial images): the problem's p, its A p (the
// normal operator's dst), the solver's partials array and the work-groups
// [blk0, blk0 + nblk) of its vector launches, len floats.
struct CgApFuse {

	float lambda;
	const float* p;
	float* Ap;
	double* part;
	int blk0;
	int nblk;
	long len;
	const int* skip;
};

// f.Ap = sum_g grp[g N + i] (groups added in order) + lambda f.p, with the
// p . Ap partials; grp holds G partial images of N complex elements
void cg_apdot_groups(const CgApFuse& f, const float2* grp, long N, int G, hipStream_t st);

// Monitor of the iterates (iter/monitor.c, pics -T): ||truth - x_k|| /
// ||truth|| of the iterate x_k that iteration k starts from, as BART's
// conjgrad calls its monitor (iter/italgos.c:534).  truth is packed like x
// (null: nothing is monitored); tinv[q] = 1 / ||truth_q|| (0: problem q is
// not monitored); part: partial sums per work-group; err: [B][tstride].
struct CgMon {

	const float* truth;
	const float* tinv;
	double* part;
	float* err;
};

// normal(i, dst, src, skip, fuse): enqueue dst = A_i^H A_i src for problem i
// on the batch's stream; skip is a device flag (may be null) that turns the
// launches into no-ops once problem i has stopped.  With fuse (may be null)
// the operator may finish with the CG step instead (dst = A^H A src +
// lambda src and the partials): it returns whether it did.
using BatchNormalFn = std::function<bool(int, float2*, const float2*, const int*, const CgApFuse*)>;

// normal(dst, src, skip, fuse) of a single problem
using NormalFn = std::function<bool(float2*, const float2*, const int*, const CgApFuse*)>;

// Partial images a grouped normal operator leaves for problem q: the CG step
// sums them (G of them, gstride floats apart; grp = A p itself when G = 1)
// and adds lambda p; grp = null: the problem is not in a group.
struct CgGrpItem {

	const float* grp;
	long gstride;
	int G;
};

struct CgBatch;

// Normal operators of many problems applied together: prepare(bt) builds
// their tables before the iteration is captured and fills items (one per
// problem); launch(bt) enqueues the grouped passes on bt.stream.  Problems
// with items[q].grp == null go through the per-problem normal function.
struct CgGroupHooks {

	std::function<void(CgBatch&, std::vector<CgGrpItem>&)> prepare;
	std::function<void(CgBatch&)> launch;
};

// Independent problems (A_i^H A_i + lambda I) x_i = b_i solved in lockstep:
// the vectors of all problems are packed into one buffer each, every vector
// operation is one launch over all problems, and the scalars (alpha, beta,
// stop flags) are per problem -- iter/italgos.c:589-680 (conjgrad_batch)
// with the per-problem stopping rule of the single solve.
struct CgBatch {

	int B = 0;
	long total = 0;    // floats over all problems
	int nblk = 0;      // work-groups of a vector launch
	hipStream_t stream = 0;
	std::vector<CgSeg> seg_h;
	DevBuf<float> x, b, r, p, Ap, trace;
	DevBuf<double> part, part_b, part2;  // p.Ap (and ||r0||^2), ||b||^2, r.r
	// monitor (single problem): truth image, 1 / ||truth||, partials, errors
	DevBuf<float> mon_truth, mon_tinv, mon_err;
	DevBuf<double> mon_part;
	bool monitored = false;
	CgMon mon() const;
	DevBuf<CgSeg> seg;
	DevBuf<int> bprob;   // work-group -> problem
	DevBuf<int> mask;    // problems to (re)initialise
	DevBuf<CgState> st;
	CgState* host_st = nullptr;
	CgState* probe = nullptr;            // stop flags read between chunks (2 x B)
	hipEvent_t probe_ev[2] = { nullptr, nullptr };
	int trace_len = 0;   // per problem
	std::vector<char> fell_back;

	// grouped normal operators (null: every problem by the normal function)
	const CgGroupHooks* hooks = nullptr;
	bool hooks_ready = false;
	std::vector<CgGrpItem> grp_h;
	DevBuf<CgGrpItem> grp;

	hipGraph_t graph = nullptr;
	hipGraphExec_t gexec = nullptr;
	bool have_graph = false;
	float graph_lambda = 0.f;

	// nfloats[i]: floats (2 x complex elements) of problem i
	CgBatch(const std::vector<long>& nfloats, hipStream_t s);
	~CgBatch();
	CgBatch(const CgBatch&) = delete;
	CgBatch& operator=(const CgBatch&) = delete;

	float* xs(int i) const { return x.ptr + seg_h[i].off; }
	float* bs(int i) const { return b.ptr + seg_h[i].off; }
	float* rs(int i) const { return r.ptr + seg_h[i].off; }

	void drop_graph();

	// truth (device, n floats of problem 0; null: off) for the monitor of
	// the next solves
	void set_monitor(const float* truth, double tnorm);
	// errors of the last solve (host, iterations entries)
	void monitor_errors(float* host, int n) const;

	// Enqueues the start: b_i copied in, x_i = warm start or zero, r = b - (A^H A
	// + lambda I) x, p = r.  With warm && fallback, waits for the device and
	// restarts from zero every problem whose warm-start residual exceeds ||b_i||.
	void start(const BatchNormalFn& normal, float lambda, float tol, int maxiter, const float2* const* x0,
		   const float2* const* b_in, bool warm, bool fallback);

	// Enqueues up to maxiter iterations: one iteration is captured once as a
	// hipGraph (per lambda) and launched in chunks until maxiter or until
	// every problem has stopped; stopped problems skip their work.
	void iterate(const BatchNormalFn& normal, float lambda, int maxiter, bool use_graph);

	// pieces of iterate_batches
	void capture(const BatchNormalFn& normal, float lambda);
	void launch(int n);
	void post_probe(int slot);
	bool probe_all_done(int slot);

	// Enqueues x_i -> x_out[i] and the state read-back.
	void finish(float2* const* x_out);

	// After the stream has completed: per-problem results.
	CgResult result(int i) const;

private:
	void enqueue_iteration(const BatchNormalFn& normal, float lambda);
	void init_kernels(float lambda, float tol, int maxiter, bool cold, bool masked);
};

// Iterations of several batches, each on its own stream, launched
// round-robin so that the streams run concurrently (see CgBatch::iterate).
void iterate_batches(const std::vector<CgBatch*>& bs, const std::vector<const BatchNormalFn*>& fns, float lambda,
		     int maxiter);

// Largest problem (complex elements) and iteration count of cg_solve_noalloc
constexpr long CG_NOALLOC_ELEMS = 1L << 21;
constexpr int CG_NOALLOC_TRACE = 4096;

// One problem solved in the library's static working memory: no device
// memory is allocated (no graph either; the iterations are launched one by
// one, in chunks after which the stop flag is read).  x (warm start in, the
// solution out) and b are device pointers of nfloats floats; trace_host (may
// be null) receives the residual norms.  Calls that overlap on one GPU race.
CgResult cg_solve_noalloc(const NormalFn& normal, float lambda, int maxiter, float tol, float2* x,
			  const float2* b, long nfloats, bool warm, bool fallback, hipStream_t stream,
			  float* trace_host);

// Single problem with its solver state kept across calls (graph reuse).
struct CgSolver {

	CgBatch bt;

	CgSolver(long nfloats, hipStream_t s) : bt(std::vector<long>{ nfloats }, s) {}

	// x_io, b: device pointers of n floats.  warm: x_io holds the initial
	// guess; fallback: restart from zero when the warm start's residual
	// exceeds ||b||.
	CgResult solve(const NormalFn& normal, float lambda, int maxiter, float tol, float2* x_io, const float2* b,
		       bool warm, bool fallback, bool use_graph);

	// with a truth image (device; null: off), the monitor of the next solves
	void set_monitor(const float2* truth, double tnorm) { bt.set_monitor((const float*)truth, tnorm); }

	void drop_graph() { bt.drop_graph(); }
};

}  // namespace bam
