// linop.hip -- the C ABI (include/bart_amd.h): BART's linop surface over the
// gfx950 NUFFT plan (linops/linop.c, noncart/nufft.c:1174-1327,
// sense/modelnc.c:28-67), FFT entry points and the CG solver.
#include "../../include/bart_amd.h"

#include "cg.hpp"
#include "cnn.hpp"
#include "espirit.hpp"
#include "nudft.hpp"
#include "nufft_impl.hpp"

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

namespace bam {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static std::atomic<int> error_mode{ BART_AMD_ERRORS_EXIT };
static thread_local std::string last_error;
static thread_local bool has_error = false;

void fail(const char* fmt, ...)
{
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);

	if (BART_AMD_ERRORS_EXIT == error_mode.load()) {
		fprintf(stderr, "Error: %s\n", buf);
		exit(1);
	}

	throw Error{ buf };
}

template <class F> static auto guarded(F&& f, decltype(f()) on_error) -> decltype(f())
{
	try {
		return f();
	} catch (const Error& e) {
		last_error = e.msg;
		has_error = true;
	} catch (const std::exception& e) {
		last_error = e.what();
		has_error = true;
		if (BART_AMD_ERRORS_EXIT == error_mode.load()) {
			fprintf(stderr, "Error: %s\n", e.what());
			exit(1);
		}
	}
	return on_error;
}

template <class F> static void guarded_void(F&& f)
{
	guarded([&]() { f(); return 0; }, 0);
}

hipStream_t default_stream() { return 0; }

// ---------------------------------------------------------------------------
// pointer location (BART: cuda_ondevice, num/gpuops.c:445)
// ---------------------------------------------------------------------------

static bool on_device(const void* p)
{
	if (nullptr == p)
		return false;

	hipPointerAttribute_t attr;
	hipError_t e = hipPointerGetAttributes(&attr, p);

	if (hipSuccess != e) {
		(void)hipGetLastError();
		return false;
	}

	return (hipMemoryTypeDevice == attr.type) || (hipMemoryTypeManaged == attr.type);
}

// ---------------------------------------------------------------------------
// operator handle
// ---------------------------------------------------------------------------

enum OpKind { OP_NUFFT = 1, OP_SENSE = 2, OP_NUDFT = 3 };

// stream and lock of one NUFFT plan, shared by every handle that applies it
struct Queue {

	hipStream_t stream = 0;
	std::mutex mtx;

	Queue() { BAM_HIP(hipStreamCreate(&stream)); }
	~Queue() { (void)hipStreamDestroy(stream); }
};

struct Op {

	int kind = 0;
	int N = 0;
	long odims[32];
	long idims[32];
	long osize = 0, isize = 0;
	std::shared_ptr<Queue> q;
	std::shared_ptr<Nufft> nufft;
	std::shared_ptr<Nudft> nudft;
	std::shared_ptr<Chain> chain;  // os != 2 (nufft_create_chain); nufft is null then
	std::shared_ptr<DevBuf<float2>> sens;
	hipStream_t stream = 0;       // the handle's own stream
	hipStream_t user_stream = 0;  // set by bart_amd_linop_set_stream (0: none)
	bool normal_only = false;     // nufft_create_normal: no forward / adjoint
	// low-memory modes (lowmem, zero_overhead): the plan holds a chunk of
	// cchunk coils (its sub-grid stack S cchunk N instead of S C N) and is
	// applied chunk by chunk; ctotal coils in all
	int ctotal = 0, cchunk = 0;
	DevBuf<float2> lm_part;       // per-chunk partial images of S^H (SENSE)
	// subspace basis (K > 0): ncoil coils, K coefficients, T frames of Mf
	// samples; the plan's channels are the coil-coefficient pairs.  bksp:
	// coefficient k-space (ncoil K T Mf), bcim: coefficient coil images of
	// a SENSE operator (ncoil K N)
	int K = 0, T = 1, ncoil = 0;
	long Mf = 0;
	// batch dims >= 4 sharing the trajectory: nbatch images; a SENSE
	// operator then applies its maps to each (bcim: the ncoil nbatch coil
	// images), the plan has ncoil nbatch channels
	int nbatch = 1;
	// frames with trajectories of their own: one plan each (nufft = the
	// first), frame t's k-space and images at t times the frame slab
	std::vector<std::shared_ptr<Nufft>> frames;
	DevBuf<float2> bksp, bcim;
	std::mutex& mtx() { return q->mtx; }
	DevBuf<float2> stage_in, stage_out, kbuf;
	std::unique_ptr<CgSolver> cg;
	float* mon_err = nullptr;     // host array of bart_amd_cg_monitor (null: off)
	// warm-start guess: A^H A b, partial sums, scalars, network channels in / out
	DevBuf<float2> g_ab, g_b, g_d, g_x, g_kw, g_dcf;
	DevBuf<float> g_w;
	DevBuf<double> g_part;
	DevBuf<float> g_sc, g_in, g_out;

	~Op()
	{
		cg.reset();
		nufft.reset();
		nudft.reset();
		q.reset();
	}

	const float2* maps() const { return (OP_SENSE == kind) ? sens->ptr : nullptr; }

	int nchunks() const { return (cchunk > 0) ? ctotal / cchunk : 1; }

	// chunk k's coil images, maps and k-space
	long cim_off(int k) const { return (long)k * cchunk * nufft->N; }
	long ksp_off(int k) const { return (long)k * cchunk * nufft->M; }

	// S^H over the chunks: partial image k of chunk k, summed in chunk order
	// (allocated with the SENSE operator, so that no allocation happens
	// inside a captured CG iteration)
	float2* chunk_part(int k)
	{
		BAM_CHECK((long)lm_part.n >= (long)nchunks() * nufft->N, "low-memory SENSE: no partial images");
		return lm_part.ptr + (long)k * nufft->N;
	}

	// device pointers, on the plan's current stream
	void fwd(float2* d, const float2* s)
	{
		BAM_CHECK(!normal_only, "NuFFT with normal operator only!");
		if (OP_NUDFT == kind) {
			nudft->forward(d, s);
			return;
		}
		if (!frames.empty()) {
			const float2* m = maps();
			const long ci = (nullptr != m) ? nufft->N : (long)ncoil * nufft->N;
			for (size_t t = 0; t < frames.size(); t++)
				frames[t]->forward(d + (long)t * ncoil * nufft->M, s + (long)t * ci, m);
			return;
		}
		if (chain) {
			const float2* m = maps();
			if (nullptr != m)
				maps_expand(bcim.ptr, m, s, ncoil, nbatch, chain->N, chain->stream);
			chain->forward(d, (nullptr != m) ? bcim.ptr : s);
			return;
		}
		if (nbatch > 1 && nullptr != maps()) {
			maps_expand(bcim.ptr, maps(), s, ncoil, nbatch, nufft->N, nufft->stream);
			nufft->forward(d, bcim.ptr, nullptr);
			return;
		}
		if (K > 0) {
			// maps, coefficient NUFFT, contraction with the basis
			// (noncart/nufft.c:1450-1453)
			const float2* m = maps();
			const float2* cim = s;
			if (nullptr != m) {
				maps_expand(bcim.ptr, m, s, ncoil, K, nufft->N, nufft->stream);
				cim = bcim.ptr;
			}
			nufft->forward(bksp.ptr, cim, nullptr);
			basis_contract(d, bksp.ptr, nufft->bas.ptr, ncoil, K, T, Mf, nufft->stream);
			return;
		}
		for (int k = 0; k < nchunks(); k++) {
			const float2* m = maps();
			nufft->forward(d + ksp_off(k), m ? s : s + cim_off(k), m ? m + cim_off(k) : nullptr);
		}
	}

	void adj(float2* d, const float2* s)
	{
		BAM_CHECK(!normal_only, "NuFFT with normal operator only!");
		if (OP_NUDFT == kind) {
			nudft->adjoint(d, s);
			return;
		}
		if (!frames.empty()) {
			const float2* m = maps();
			const long ci = (nullptr != m) ? nufft->N : (long)ncoil * nufft->N;
			for (size_t t = 0; t < frames.size(); t++)
				frames[t]->adjoint(d + (long)t * ci, s + (long)t * ncoil * nufft->M, m);
			return;
		}
		if (chain) {
			const float2* m = maps();
			chain->adjoint((nullptr != m) ? bcim.ptr : d, s);
			if (nullptr != m)
				maps_reduce(d, m, bcim.ptr, ncoil, nbatch, chain->N, chain->stream);
			return;
		}
		if (nbatch > 1 && nullptr != maps()) {
			nufft->adjoint(bcim.ptr, s, nullptr);
			maps_reduce(d, maps(), bcim.ptr, ncoil, nbatch, nufft->N, nufft->stream);
			return;
		}
		if (K > 0) {
			// x conj(basis) (noncart/nufft.c:1483-1487), coefficient
			// adjoint NUFFT, maps
			const float2* m = maps();
			basis_expand(bksp.ptr, s, nufft->bas.ptr, ncoil, K, T, Mf, nufft->stream);
			nufft->adjoint((nullptr != m) ? bcim.ptr : d, bksp.ptr, nullptr);
			if (nullptr != m)
				maps_reduce(d, m, bcim.ptr, ncoil, K, nufft->N, nufft->stream);
			return;
		}
		const float2* m = maps();
		if (1 == nchunks() || nullptr == m) {
			for (int k = 0; k < nchunks(); k++)
				nufft->adjoint(d + (m ? 0 : cim_off(k)), s + ksp_off(k), m);
			return;
		}
		for (int k = 0; k < nchunks(); k++)
			nufft->adjoint(chunk_part(k), s + ksp_off(k), m + cim_off(k));
		sum_images(d, lm_part.ptr, nufft->N, nchunks(), nufft->stream, nullptr);
	}

	bool toeplitz() const { return (OP_NUDFT != kind) && nullptr != nufft && nufft->toeplitz; }

	// A^H A: Toeplitz when available, else adjoint(forward(x)) through kbuf
	// (linops/linop.c:489-525); kbuf must be allocated (prepare_normal)
	// (with fuse: the CG step after it may be fused into the operator's last
	// kernel; returns whether it was)
	bool nrm(float2* d, const float2* s, const int* skip, const CgApFuse* fuse = nullptr)
	{
		if (!toeplitz()) {
			fwd(kbuf.ptr, s);
			adj(d, kbuf.ptr);
			return false;
		}

		if (!frames.empty()) {
			const float2* m = maps();
			const long ci = (nullptr != m) ? nufft->N : (long)ncoil * nufft->N;
			for (size_t t = 0; t < frames.size(); t++)
				frames[t]->normal(d + (long)t * ci, s + (long)t * ci, m, skip);
			return false;
		}

		if (nbatch > 1 && nullptr != maps()) {
			maps_expand(bcim.ptr, maps(), s, ncoil, nbatch, nufft->N, nufft->stream, skip);
			nufft->normal(bcim.ptr, bcim.ptr, nullptr, skip);
			maps_reduce(d, maps(), bcim.ptr, ncoil, nbatch, nufft->N, nufft->stream, skip);
			return false;
		}

		if (K > 0) {
			const float2* m = maps();
			if (nullptr == m) {
				nufft->normal_basis(d, s, skip);
				return false;
			}
			maps_expand(bcim.ptr, m, s, ncoil, K, nufft->N, nufft->stream, skip);
			nufft->normal_basis(bcim.ptr, bcim.ptr, skip);
			maps_reduce(d, m, bcim.ptr, ncoil, K, nufft->N, nufft->stream, skip);
			return false;
		}

		if (1 == nchunks())
			return nufft->normal(d, s, maps(), skip, fuse);

		const float2* m = maps();

		if (nullptr == m) {
			for (int k = 0; k < nchunks(); k++)
				nufft->normal(d + cim_off(k), s + cim_off(k), nullptr, skip);
			return false;
		}

		for (int k = 0; k < nchunks(); k++)
			nufft->normal(chunk_part(k), s, m + cim_off(k), skip);

		if (nullptr != fuse) {
			cg_apdot_groups(*fuse, lm_part.ptr, nufft->N, nchunks(), nufft->stream);
			return true;
		}

		sum_images(d, lm_part.ptr, nufft->N, nchunks(), nufft->stream, skip);
		return false;
	}

	void prepare_normal()
	{
		if (!toeplitz() && (long)kbuf.n < osize)
			kbuf.alloc(osize);
	}

	// the plan's launches go to stream s (the operator's own stream by default)
	void set_plan_stream(hipStream_t s)
	{
		if (nufft)
			nufft->stream = s;
		if (chain)
			chain->stream = s;
		for (auto& f : frames)
			f->stream = s;
		if (nudft)
			nudft->stream = s;
	}

	const void* plan_id() const
	{
		return nufft ? (const void*)nufft.get() : (chain ? (const void*)chain.get() : (const void*)nudft.get());
	}

	// relative cost of one normal operator: the sub-grid stack it transforms
	double cost() const
	{
		if (OP_NUDFT == kind)
			return (double)isize * (double)osize;
		if (chain)
			return (double)chain->C * (double)chain->N2;
		if (!frames.empty())
			return (double)frames.size() * nufft->S * nufft->C * (double)nufft->N;
		return (double)nufft->S * nufft->C * nchunks() * (double)nufft->N;
	}

	Nufft& plan(const char* what)
	{
		BAM_CHECK(OP_NUDFT != kind && (nullptr != nufft || nullptr != chain), "%s: not a NUFFT operator", what);
		BAM_CHECK(nullptr == chain, "%s: not supported for the chained NUFFT (oversampling other than 2)", what);
		BAM_CHECK(frames.empty(), "%s: not supported for per-frame trajectories", what);
		BAM_CHECK(0 == K, "%s: not supported for an operator with a subspace basis", what);
		return *nufft;
	}
};

static long md_size(int N, const long* d)
{
	long s = 1;
	for (int i = 0; i < N; i++)
		s *= d[i];
	return s;
}

static void check_dims(const char* what, int N, const long* a, int M, const long* b)
{
	BAM_CHECK(N == M, "%s: dimension count %d != %d", what, N, M);

	for (int i = 0; i < N; i++)
		BAM_CHECK(a[i] == b[i], "%s: dims differ in dim %d (%ld != %ld)", what, i, a[i], b[i]);
}

// copy of a host or device array into host memory
template <class T> static std::vector<T> to_host(const T* p, long n)
{
	std::vector<T> h(n);

	if (n > 0) {
		if (on_device(p))
			BAM_HIP(hipMemcpy(h.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost));
		else
			memcpy(h.data(), p, sizeof(T) * n);
	}

	return h;
}

// Applies f(dst_dev, src_dev) with staging of host pointers and a final sync.
// The launches go to the handle's own stream, or to the stream set with
// bart_amd_linop_set_stream; with device pointers on a stream of the
// caller's the call returns without waiting (the caller orders its work on
// that stream), otherwise it returns when the result is in dst.
template <class F>
static void apply_staged(Op* op, float2* dst, long dn, const float2* src, long sn, F&& f)
{
	std::lock_guard<std::mutex> lock(op->mtx());

	const hipStream_t st = (0 != op->user_stream) ? op->user_stream : op->stream;
	op->set_plan_stream(st);

	bool dd = on_device(dst);
	bool sd = on_device(src);

	const float2* s = src;
	float2* d = dst;

	if (!sd) {
		if ((long)op->stage_in.n < sn)
			op->stage_in.alloc(sn);
		BAM_HIP(hipMemcpyAsync(op->stage_in.ptr, src, sizeof(float2) * sn, hipMemcpyHostToDevice, st));
		s = op->stage_in.ptr;
	}

	if (!dd) {
		if ((long)op->stage_out.n < dn)
			op->stage_out.alloc(dn);
		d = op->stage_out.ptr;
	}

	f(d, s);

	if (!dd)
		BAM_HIP(hipMemcpyAsync(dst, d, sizeof(float2) * dn, hipMemcpyDeviceToHost, st));

	if (0 == op->user_stream || !dd || !sd)
		BAM_HIP(hipStreamSynchronize(st));

	op->set_plan_stream(op->stream);
}

// Parse BART dims for a NUFFT: image dims 0-2, coils dim 3, samples dims 1-2
// of ksp/traj (noncart/nufft.c:1208-1238 checks).
struct NufftDims {
	int img[3];
	int C;
	long M;
	int B;  // batch: the product of dims >= 4 (trajectory and weights shared)
};

static NufftDims parse_nufft_dims(int N, const long* ksp_dims, const long* cim_dims, const long* traj_dims)
{
	BAM_CHECK(N >= 4 && N <= 32, "nufft: unsupported dimension count %d", N);
	BAM_CHECK(3 == traj_dims[0], "nufft: trajectory dim 0 must be 3 (got %ld)", traj_dims[0]);
	BAM_CHECK(1 == ksp_dims[0], "nufft: k-space dim 0 must be 1");
	BAM_CHECK(ksp_dims[1] == traj_dims[1] && ksp_dims[2] == traj_dims[2],
		  "Incompatible dimensions of k-space and trajectory!");
	BAM_CHECK(1 == traj_dims[3], "nufft: trajectory must not have a coil dim");
	BAM_CHECK(ksp_dims[3] == cim_dims[3], "nufft: coil dims of k-space (%ld) and image (%ld) differ",
		  ksp_dims[3], cim_dims[3]);

	// dims >= 4 are batch dims with one trajectory for all (the NUFFT then
	// treats them like coils: k-space and images match there,
	// noncart/nufft.c:1229)
	long B = 1;

	for (int i = 4; i < N; i++) {
		BAM_CHECK(ksp_dims[i] == cim_dims[i], "nufft: k-space and image differ in dim %d (%ld != %ld)", i,
			  ksp_dims[i], cim_dims[i]);
		BAM_CHECK(1 == traj_dims[i], "nufft: the trajectory may not vary over dim %d (one trajectory per batch)", i);
		B *= ksp_dims[i];
	}

	BAM_CHECK(B >= 1 && B < (1l << 24), "nufft: bad batch size %ld", B);

	NufftDims r;
	r.B = (int)B;
	for (int i = 0; i < 3; i++) {
		BAM_CHECK(cim_dims[i] >= 1 && cim_dims[i] < (1l << 20), "nufft: bad image dim");
		r.img[i] = (int)cim_dims[i];
	}
	r.C = (int)cim_dims[3];
	r.M = ksp_dims[1] * ksp_dims[2];
	BAM_CHECK(r.C >= 1, "nufft: no coils");
	return r;
}

// upper_triag (noncart/nufft.c:1070-1074) stores the Hermitian K x K PSF
// as its upper triangle; the full K x K storage kept here holds the same
// values, so the option is accepted
static void check_conf(const nufft_conf_s& conf, bool has_basis)
{
	(void)has_basis;
	BAM_CHECK(!conf.pcycle, "nufft: phase cycling is not supported");
	BAM_CHECK(!conf.compress_psf, "nufft: compressed PSF storage is not supported");
	BAM_CHECK(0 == conf.cfft, "nufft: cfft flags are not supported");
	BAM_CHECK(7 == (conf.flags & 7), "nufft: flags must include dims 0-2");
}

// trajectory [3][M] (complex, host or device) -> host M x 3 reals
static std::vector<float> traj_reals_of(const float2* traj, long M)
{
	std::vector<float2> th = to_host(traj, 3 * M);
	std::vector<float> tr(3 * M);

	for (long i = 0; i < 3 * M; i++)
		tr[i] = th[i].x;

	return tr;
}

// weights of shape wgh_dims (broadcast over coils) -> per-sample vector;
// with T > 1 frames (subspace basis) the weights may vary over dim 5 and
// the vector holds the frames' samples r + R (s + Sp t).  Over batch dims
// (>= 4, sharing the trajectory) the weights must be the same for every
// batch entry (as the pattern pics estimates from repeated data is); they
// may be given broadcast or in full.
static std::vector<float2> expand_weights(int N, const long* wgh_dims, const float2* w, const long* ksp_dims,
					  long T = 1)
{
	std::vector<float2> out;

	if (nullptr == w)
		return out;

	long R = ksp_dims[1], Sp = ksp_dims[2];
	long wr = wgh_dims[1], ws = wgh_dims[2];
	long wt = (N > 5 && T > 1) ? wgh_dims[5] : 1;

	BAM_CHECK(1 == wgh_dims[0], "nufft: weights dim 0 must be 1");
	BAM_CHECK((wr == R || wr == 1) && (ws == Sp || ws == 1), "Incompatible dimensions of k-space and weights!");
	BAM_CHECK(wt == 1 || wt == T, "Incompatible dimensions of k-space and weights!");
	BAM_CHECK(1 == wgh_dims[3], "nufft: weights must not vary over the coil dim");

	long str[32];
	long tot = 1;

	for (int i = 0; i < N; i++) {
		str[i] = tot;
		tot *= wgh_dims[i];
	}

	for (int i = 4; i < N; i++)
		BAM_CHECK(1 == wgh_dims[i] || wgh_dims[i] == ksp_dims[i], "Incompatible dimensions of k-space and weights!");

	std::vector<float2> h = to_host(w, tot);
	out.resize(R * Sp * T);

	for (long t = 0; t < T; t++)
		for (long s = 0; s < Sp; s++)
			for (long r = 0; r < R; r++)
				out[r + R * (s + Sp * t)] =
					h[(wr > 1 ? r : 0) * str[1] + (ws > 1 ? s : 0) * str[2] + (wt > 1 ? t * str[5] : 0)];

	// batch dims: every batch entry must hold the same weights
	for (long e = 0; e < tot; e++) {
		long rem = e, base = 0;
		for (int i = 0; i < N; i++) {
			const long c = rem % wgh_dims[i];
			rem /= wgh_dims[i];
			if (i < 4 || (5 == i && T > 1))
				base += c * str[i];
		}
		BAM_CHECK(h[e].x == h[base].x && h[e].y == h[base].y,
			  "nufft: weights that differ between batch entries (dims >= 4) are not supported");
	}

	return out;
}

// NUFFT with a subspace basis (noncart/nufft.c:1018-1036, 1208-1258):
// basis [1 1 1 1 1 T K], k-space [1 R S C 1 T (1|K)] (the operator's
// output [1 R S C 1 T 1]), coil images [X Y Z C 1 1 K], trajectory
// [3 R S 1 1 (1|T)], weights [1 (1|R) (1|S) 1 1 (1|T)].  The plan grids the
// C K coefficient images at the T Mf samples of all frames.
static Op* make_basis_op(int N, const long* ksp_dims, const long* cim_dims, const long* traj_dims, const float2* traj,
			 const long* wgh_dims, const float2* weights, const nufft_conf_s& conf, const long* bas_dims,
			 const float2* basis)
{
	BAM_CHECK(N >= 7 && N <= 32, "nufft: a subspace basis needs at least 7 dims (got %d)", N);

	for (int i = 0; i < N; i++)
		BAM_CHECK(1 == bas_dims[i] || 5 == i || 6 == i, "nufft: the basis may vary over dims 5 and 6 only");

	const long T = bas_dims[5], K = bas_dims[6];
	BAM_CHECK(K >= 1 && K <= BASIS_MAX, "nufft: %ld basis coefficients (at most %d)", K, BASIS_MAX);
	BAM_CHECK(T >= 1, "nufft: bad basis dims");
	BAM_CHECK(ksp_dims[5] == T, "nufft: k-space dim 5 (%ld) does not match the basis (%ld)", ksp_dims[5], T);
	BAM_CHECK(1 == ksp_dims[6] || K == ksp_dims[6], "nufft: k-space dim 6 must be 1 or %ld", K);
	BAM_CHECK(1 == cim_dims[5] && K == cim_dims[6],
		  "nufft: coil images must have size 1 in dim 5 and %ld (coefficients) in dim 6", K);
	BAM_CHECK(1 == traj_dims[6] && (1 == traj_dims[5] || T == traj_dims[5]),
		  "nufft: trajectory dim 5 must be 1 or %ld, dim 6 must be 1", T);

	long k1[32], c1[32], t1[32], out[32];

	for (int i = 0; i < N; i++) {
		k1[i] = (5 == i || 6 == i) ? 1 : ksp_dims[i];
		c1[i] = (6 == i) ? 1 : cim_dims[i];
		t1[i] = (5 == i) ? 1 : traj_dims[i];
		out[i] = (6 == i) ? 1 : ksp_dims[i];
	}

	NufftDims nd = parse_nufft_dims(N, k1, c1, t1);
	BAM_CHECK(1 == nd.B, "nufft: a subspace basis with further batch dims is not supported");
	const long Mf = nd.M;
	const long Mp = Mf * T;
	const bool traj_t = (T > 1 && T == traj_dims[5]);

	std::vector<float2> th = to_host(traj, 3 * Mf * (traj_t ? T : 1));
	std::vector<float> tr(3 * Mp);

	for (long t = 0; t < T; t++)
		for (long m = 0; m < Mf; m++)
			for (int d = 0; d < 3; d++)
				tr[3 * (m + Mf * t) + d] = th[3 * (m + Mf * (traj_t ? t : 0)) + d].x;

	std::vector<float2> w = expand_weights(N, wgh_dims, weights, ksp_dims, T);
	std::vector<float2> B = to_host(basis, K * T);

	auto op = new Op();
	op->kind = OP_NUFFT;
	op->N = N;
	memcpy(op->odims, out, sizeof(long) * N);
	memcpy(op->idims, cim_dims, sizeof(long) * N);
	op->osize = md_size(N, out);
	op->isize = md_size(N, cim_dims);
	op->q = std::make_shared<Queue>();
	op->stream = op->q->stream;
	op->K = (int)K;
	op->T = (int)T;
	op->Mf = Mf;
	op->ncoil = nd.C;
	// the low-memory modes give the same results; with a basis the plan
	// keeps all coils (no chunks)
	op->ctotal = nd.C * (int)K;
	op->cchunk = op->ctotal;

	try {
		op->nufft = std::make_shared<Nufft>();
		op->nufft->stream = op->stream;
		op->nufft->real_psf = conf.real;
		op->nufft->set_basis((int)K, (int)T, Mf, B.data());
		op->nufft->create(nd.img, nd.C * (int)K, Mp, tr.data(), w.empty() ? nullptr : w.data(), conf.toeplitz,
				  conf.periodic, conf.width, conf.os, conf.toeplitz && !conf.nopsf);
		op->bksp.alloc((size_t)nd.C * K * Mp);
	} catch (...) {
		delete op;
		throw;
	}

	return op;
}

// Batch dims with a trajectory per entry (frames, noncart/nufft.c:1229 with
// trajectory dims > 1 beyond dim 3): one plan per frame, applied in frame
// order to the frame's slab of k-space and images; the weights may vary
// over the frames as well.
static Op* make_frames_op(int N, const long* ksp_dims, const long* cim_dims, const long* traj_dims,
			  const float2* traj, const long* wgh_dims, const float2* weights, const nufft_conf_s& conf,
			  long F)
{
	long t1[32], w1[32];
	long wstride = 0;
	bool wframes = false;

	for (int i = 0; i < N; i++) {
		t1[i] = (i >= 4) ? 1 : traj_dims[i];
		w1[i] = (nullptr != wgh_dims && i >= 4) ? 1 : (nullptr != wgh_dims ? wgh_dims[i] : 1);
		if (nullptr != wgh_dims && i >= 4 && wgh_dims[i] > 1) {
			BAM_CHECK(wgh_dims[i] == ksp_dims[i] && traj_dims[i] == ksp_dims[i],
				  "Incompatible dimensions of k-space and weights!");
			wframes = true;
		}
	}

	NufftDims nd = parse_nufft_dims(N, ksp_dims, cim_dims, t1);
	BAM_CHECK(nd.B == F, "nufft: frames and batch dims do not match");

	if (wframes) {
		wstride = 1;
		for (int i = 0; i < 4; i++)
			wstride *= wgh_dims[i];
	}

	auto op = new Op();
	op->kind = OP_NUFFT;
	op->N = N;
	memcpy(op->odims, ksp_dims, sizeof(long) * N);
	memcpy(op->idims, cim_dims, sizeof(long) * N);
	op->osize = md_size(N, ksp_dims);
	op->isize = md_size(N, cim_dims);
	op->q = std::make_shared<Queue>();
	op->stream = op->q->stream;
	op->ctotal = op->cchunk = nd.C;
	op->ncoil = nd.C;
	op->nbatch = (int)F;

	try {
		for (long t = 0; t < F; t++) {
			std::vector<float> tr = traj_reals_of(traj + 3 * nd.M * t, nd.M);
			std::vector<float2> w = expand_weights(N, w1, weights ? weights + wstride * t : nullptr, ksp_dims);
			auto p = std::make_shared<Nufft>();
			p->stream = op->stream;
			p->real_psf = conf.real;
			p->create(nd.img, nd.C, nd.M, tr.data(), w.empty() ? nullptr : w.data(), conf.toeplitz,
				  conf.periodic, conf.width, conf.os, conf.toeplitz && !conf.nopsf);
			op->frames.push_back(p);
		}
	} catch (...) {
		delete op;
		throw;
	}

	op->nufft = op->frames[0];
	return op;
}

static Op* make_nufft_op(int N, const long* ksp_dims, const long* cim_dims, const long* traj_dims, const float2* traj,
			 const long* wgh_dims, const float2* weights, const nufft_conf_s& conf,
			 const long* bas_dims = nullptr, const float2* basis = nullptr)
{
	if (nullptr != basis) {
		BAM_CHECK(2.f == conf.os && conf.decomp,
			  "nufft: a subspace basis needs the decomposed NUFFT with oversampling 2");
		BAM_CHECK(nullptr != bas_dims, "nufft: basis without dims");
		return make_basis_op(N, ksp_dims, cim_dims, traj_dims, traj, wgh_dims, weights, conf, bas_dims, basis);
	}

	// batch dims over which the trajectory varies (frames with trajectories
	// of their own): one plan per frame
	long F = 1;
	bool shared = false;

	for (int i = 4; i < N; i++) {
		if (traj_dims[i] > 1) {
			BAM_CHECK(traj_dims[i] == ksp_dims[i], "Incompatible dimensions of k-space and trajectory!");
			F *= traj_dims[i];
		} else if (ksp_dims[i] > 1) {
			shared = true;
		}
	}

	if (F > 1) {
		BAM_CHECK(!shared, "nufft: batch dims with and without trajectories of their own cannot be mixed");
		BAM_CHECK(2.f == conf.os && conf.decomp, "nufft: per-frame trajectories need the decomposed NUFFT with oversampling 2");
		return make_frames_op(N, ksp_dims, cim_dims, traj_dims, traj, wgh_dims, weights, conf, F);
	}

	NufftDims nd = parse_nufft_dims(N, ksp_dims, cim_dims, traj_dims);

	if (2.f != conf.os || !conf.decomp) {

		// the chained NUFFT (noncart/nufft.c:1184-1199), or the
		// non-decomposed one (decomp = false, :851, 898-935: no Toeplitz
		// mode); no memory modes
		BAM_CHECK(2.f != conf.os || !conf.toeplitz,
			  "nufft: the non-decomposed grid (-1) has no Toeplitz operator (nufft.c:851)");
		std::vector<float> tr = traj_reals_of(traj, nd.M);
		std::vector<float2> w = expand_weights(N, wgh_dims, weights, ksp_dims);
		auto op = new Op();
		op->kind = OP_NUFFT;
		op->N = N;
		memcpy(op->odims, ksp_dims, sizeof(long) * N);
		memcpy(op->idims, cim_dims, sizeof(long) * N);
		op->osize = md_size(N, ksp_dims);
		op->isize = md_size(N, cim_dims);
		op->q = std::make_shared<Queue>();
		op->stream = op->q->stream;
		op->ctotal = op->cchunk = nd.C * nd.B;
		op->ncoil = nd.C;
		op->nbatch = nd.B;

		try {
			op->chain = std::make_shared<Chain>();
			op->chain->stream = op->stream;
			op->chain->create(nd.img, nd.C * nd.B, nd.M, tr.data(), w.empty() ? nullptr : w.data(),
					  conf.periodic, conf.width, conf.os, 2.f != conf.os);
		} catch (...) {
			delete op;
			throw;
		}

		return op;
	}


	std::vector<float2> th = to_host(traj, 3 * nd.M);
	std::vector<float> tr(3 * nd.M);

	for (long i = 0; i < 3 * nd.M; i++)
		tr[i] = th[i].x;

	std::vector<float2> w = expand_weights(N, wgh_dims, weights, ksp_dims);

	auto op = new Op();
	op->kind = OP_NUFFT;
	op->N = N;
	memcpy(op->odims, ksp_dims, sizeof(long) * N);
	memcpy(op->idims, cim_dims, sizeof(long) * N);
	op->osize = md_size(N, ksp_dims);
	op->isize = md_size(N, cim_dims);
	op->q = std::make_shared<Queue>();
	op->stream = op->q->stream;

	// low-memory modes (noncart/nufft.c:1629-2003 keeps one sub-grid at a
	// time, C N grid points): here coil chunks whose stack S c N is at most
	// C N for lowmem, and single coils (S N) for zero_overhead; the largest
	// chunk size that divides C
	op->ctotal = nd.C * nd.B;
	op->cchunk = op->ctotal;
	op->ncoil = nd.C;
	op->nbatch = nd.B;

	// (with batch dims the plan keeps all channels: the chunks are of coils)
	if ((conf.lowmem || conf.zero_overhead) && 1 == nd.B) {
		int S = 1;
		for (int i = 0; i < 3; i++)
			S *= (nd.img[i] > 1) ? 2 : 1;
		int target = conf.zero_overhead ? 1 : std::max(1, nd.C / S);
		while (nd.C % target)
			target--;
		op->cchunk = target;
	}

	try {
		op->nufft = std::make_shared<Nufft>();
		op->nufft->stream = op->stream;
		op->nufft->real_psf = conf.real;
		op->nufft->create(nd.img, op->cchunk, nd.M, tr.data(), w.empty() ? nullptr : w.data(), conf.toeplitz,
				  conf.periodic, conf.width, conf.os, conf.toeplitz && !conf.nopsf);
	} catch (...) {
		delete op;
		throw;
	}

	return op;
}

// trajectory [3][M] (host or device, complex; coordinates in the real parts)
// -> host M x 3 floats
static std::vector<float> traj_reals(const float2* traj, long M)
{
	std::vector<float2> th = to_host(traj, 3 * M);
	std::vector<float> tr(3 * M);

	for (long i = 0; i < 3 * M; i++)
		tr[i] = th[i].x;

	return tr;
}

// N-d strided copy of complex floats (BART's md_copy2, strides in bytes)
struct StrideND {
	int nd;
	long dims[33];
	long dstr[33];
	long sstr[33];
};

__global__ void k_copy_strided(StrideND g, long total, char* dst, const char* src)
{
	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
		long r = e, od = 0, os = 0;
		for (int i = 0; i < g.nd; i++) {
			long q = r % g.dims[i];
			r /= g.dims[i];
			od += q * g.dstr[i];
			os += q * g.sstr[i];
		}
		*(float2*)(dst + od) = *(const float2*)(src + os);
	}
}

// copy between a dense device array and an array of the given byte strides
// (host or device); to_strided: dense -> strided, else strided -> dense
static void copy_strides(int N, const long* dims, const long* strs, float2* strided, float2* dense, bool to_strided,
			 hipStream_t st)
{
	BAM_CHECK(N >= 1 && N <= 33, "strided copy: %d dims", N);
	long total = md_size(N, dims);
	long dstr[33];
	long d = sizeof(float2);
	bool contig = true;

	for (int i = 0; i < N; i++) {
		dstr[i] = d;
		contig = contig && (1 == dims[i] || strs[i] == d);
		d *= dims[i];
	}

	const bool dev = on_device(strided);

	if (contig) {
		if (to_strided)
			BAM_HIP(hipMemcpyAsync(strided, dense, sizeof(float2) * total,
					       dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
		else
			BAM_HIP(hipMemcpyAsync(dense, strided, sizeof(float2) * total,
					       dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
		BAM_HIP(hipStreamSynchronize(st));
		return;
	}

	if (!dev) {
		std::vector<float2> h(total);
		if (!to_strided)
			for (long e = 0; e < total; e++) {
				long r = e, o = 0;
				for (int i = 0; i < N; i++) {
					o += (r % dims[i]) * strs[i];
					r /= dims[i];
				}
				h[e] = *(const float2*)((const char*)strided + o);
			}
		if (to_strided) {
			BAM_HIP(hipMemcpyAsync(h.data(), dense, sizeof(float2) * total, hipMemcpyDeviceToHost, st));
			BAM_HIP(hipStreamSynchronize(st));
			for (long e = 0; e < total; e++) {
				long r = e, o = 0;
				for (int i = 0; i < N; i++) {
					o += (r % dims[i]) * strs[i];
					r /= dims[i];
				}
				*(float2*)((char*)strided + o) = h[e];
			}
		} else {
			BAM_HIP(hipMemcpyAsync(dense, h.data(), sizeof(float2) * total, hipMemcpyHostToDevice, st));
			BAM_HIP(hipStreamSynchronize(st));
		}
		return;
	}

	StrideND g;
	g.nd = N;
	for (int i = 0; i < N; i++) {
		g.dims[i] = dims[i];
		g.dstr[i] = to_strided ? strs[i] : dstr[i];
		g.sstr[i] = to_strided ? dstr[i] : strs[i];
	}
	unsigned blocks = (unsigned)std::min<long>((total + 255) / 256, 8192);
	hipLaunchKernelGGL(k_copy_strided, dim3(std::max(1u, blocks)), dim3(256), 0, st, g, total,
			   (char*)(to_strided ? strided : dense), (const char*)(to_strided ? dense : strided));
	BAM_HIP(hipGetLastError());
	BAM_HIP(hipStreamSynchronize(st));
}

// PSF dims of a NUFFT handle: image dims, 1..., and the decomposition index
static void psf_dims_of(const Op* op, long* d)
{
	for (int i = 0; i < op->N; i++)
		d[i] = (i < 3) ? op->nufft->n[i] : 1;
	d[op->N] = op->nufft->S;
}

}  // namespace bam

using namespace bam;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

struct nufft_conf_s nufft_conf_defaults = {
	.toeplitz = true,
	.pcycle = false,
	.periodic = false,
	.lowmem = false,
	.flags = 7,
	.cfft = 0u,
	.decomp = true,
	.nopsf = false,
	.upper_triag = false,
	.real = false,
	.compress_psf = false,
	.decomposed_psf = false,
	.precomp = true,
	.precomp_linphase = true,
	.precomp_fftmod = true,
	.precomp_roll = true,
	.zero_overhead = false,
	.width = 6,
	.os = 2.,
};

void bart_amd_set_error_mode(int mode) { error_mode.store(mode); }

const char* bart_amd_last_error(void) { return has_error ? last_error.c_str() : nullptr; }

void bart_amd_clear_error(void)
{
	has_error = false;
	last_error.clear();
}

const char* bart_amd_version(void) { return "bart_amd 0.1 (gfx950)"; }

struct linop_s* nufft_create2(int N, const long ksp_dims[], const long cim_dims[], const long traj_dims[],
			      const float _Complex* traj, const long wgh_dims[], const float _Complex* weights,
			      const long bas_dims[], const float _Complex* basis, struct nufft_conf_s conf)
{
	return guarded([&]() -> struct linop_s* {
		check_conf(conf, nullptr != basis);
		BAM_CHECK(nullptr != traj, "nufft: trajectory is NULL");
		Op* op = make_nufft_op(N, ksp_dims, cim_dims, traj_dims, (const float2*)traj, wgh_dims,
				       (const float2*)weights, conf, bas_dims, (const float2*)basis);
		return (struct linop_s*)op;
	}, nullptr);
}

struct linop_s* nufft_create(int N, const long ksp_dims[], const long cim_dims[], const long traj_dims[],
			     const float _Complex* traj, const float _Complex* weights, struct nufft_conf_s conf)
{
	long wgh_dims[32];

	for (int i = 0; i < N && i < 32; i++)
		wgh_dims[i] = (0 == i) ? 1 : traj_dims[i];

	return nufft_create2(N, ksp_dims, cim_dims, traj_dims, traj, wgh_dims, weights, nullptr, nullptr, conf);
}

const struct linop_s* nudft_create(int N, unsigned long flags, const long odims[], const long idims[],
				   const long tdims[], const float _Complex* traj, const long fmdims[],
				   const float _Complex* fieldmap, const long tmdims[], const float _Complex* timemap)
{
	(void)fmdims;
	(void)tmdims;

	return guarded([&]() -> const struct linop_s* {
		BAM_CHECK(7 == flags, "nudft: only flags = 7 (dims 0-2) are supported");
		BAM_CHECK(nullptr == fieldmap && nullptr == timemap, "nudft: the B0 term is not supported");
		BAM_CHECK(nullptr != traj, "nudft: trajectory is NULL");

		NufftDims nd = parse_nufft_dims(N, odims, idims, tdims);
		std::vector<float2> th = to_host((const float2*)traj, 3 * nd.M);
		std::vector<float> tr(3 * nd.M);

		for (long i = 0; i < 3 * nd.M; i++)
			tr[i] = th[i].x;

		auto op = new Op();
		op->kind = OP_NUDFT;
		op->N = N;
		memcpy(op->odims, odims, sizeof(long) * N);
		memcpy(op->idims, idims, sizeof(long) * N);
		op->osize = md_size(N, odims);
		op->isize = md_size(N, idims);
		op->q = std::make_shared<Queue>();
		op->stream = op->q->stream;

		try {
			op->nudft = std::make_shared<Nudft>();
			op->nudft->stream = op->stream;
			op->nudft->create(nd.img, nd.C * nd.B, nd.M, tr.data());
		} catch (...) {
			delete op;
			throw;
		}

		return (const struct linop_s*)op;
	}, nullptr);
}

const struct linop_s* sense_nc_init(const long max_dims[16], const long map_dims[16], const float _Complex* maps,
				    const long ksp_dims[16], const long traj_dims[16], const float _Complex* traj,
				    const struct nufft_conf_s* _conf, const long wgs_dims[16],
				    const float _Complex* weights, const long basis_dims[16],
				    const float _Complex* basis, const struct linop_s** fft_opp,
				    unsigned long shared_img_dims)
{
	return guarded([&]() -> const struct linop_s* {
		const int N = 16;
		nufft_conf_s conf = *_conf;
		check_conf(conf, nullptr != basis);
		BAM_CHECK(0 == shared_img_dims, "sense_nc_init: shared image dims are not supported");
		BAM_CHECK(1 == max_dims[4] && 1 == map_dims[4], "sense_nc_init: only one set of maps is supported");

		long coilim_dims[16], img_dims[16];

		for (int i = 0; i < N; i++) {
			coilim_dims[i] = (4 == i) ? 1 : max_dims[i];
			img_dims[i] = (3 == i || 4 == i) ? 1 : max_dims[i];
		}

		for (int i = 0; i < 4; i++)
			BAM_CHECK(map_dims[i] == coilim_dims[i], "sense_nc_init: map dims do not match (dim %d)", i);

		for (int i = 4; i < N; i++)
			BAM_CHECK(1 == map_dims[i], "sense_nc_init: maps varying over dim %d are not supported", i);

		// the coefficient dim of the k-space follows the image
		// (sense/modelnc.c:41-43)
		long ksp2[16];
		memcpy(ksp2, ksp_dims, sizeof(ksp2));
		ksp2[6] = max_dims[6];

		Op* op = make_nufft_op(N, ksp2, coilim_dims, traj_dims, (const float2*)traj, wgs_dims,
				       (const float2*)weights, conf, basis_dims, (const float2*)basis);
		op->kind = OP_SENSE;
		memcpy(op->idims, img_dims, sizeof(long) * N);
		op->isize = md_size(N, img_dims);

		if (op->chain || (op->nbatch > 1 && op->frames.empty())) {
			try {
				op->bcim.alloc((size_t)op->ncoil * op->nbatch * (op->chain ? op->chain->N : op->nufft->N));
			} catch (...) {
				delete op;
				throw;
			}
		}

		if (op->K > 0) {
			try {
				op->bcim.alloc((size_t)op->ncoil * op->K * op->nufft->N);
			} catch (...) {
				delete op;
				throw;
			}
		}

		if (op->nchunks() > 1) {
			try {
				op->lm_part.alloc((size_t)op->nchunks() * op->nufft->N);
			} catch (...) {
				delete op;
				throw;
			}
		}

		long nmap = md_size(4, map_dims);

		try {
			op->sens = std::make_shared<DevBuf<float2>>(nmap);

			if (on_device(maps))
				BAM_HIP(hipMemcpy(op->sens->ptr, maps, sizeof(float2) * nmap, hipMemcpyDeviceToDevice));
			else
				BAM_HIP(hipMemcpy(op->sens->ptr, maps, sizeof(float2) * nmap, hipMemcpyHostToDevice));
		} catch (...) {
			delete op;
			throw;
		}

		if (nullptr != fft_opp) {

			// the NUFFT alone (coil images <-> k-space), sharing the plan,
			// stream and lock of the SENSE operator; freed separately
			Op* nop = new Op();
			nop->kind = OP_NUFFT;
			nop->N = N;
			memcpy(nop->odims, op->odims, sizeof(long) * N);
			memcpy(nop->idims, coilim_dims, sizeof(long) * N);
			nop->osize = op->osize;
			nop->isize = md_size(N, coilim_dims);
			nop->q = op->q;
			nop->stream = op->stream;
			nop->nufft = op->nufft;
			nop->chain = op->chain;
			nop->ctotal = op->ctotal;
			nop->cchunk = op->cchunk;
			nop->K = op->K;
			nop->nbatch = op->nbatch;
			nop->frames = op->frames;
			nop->T = op->T;
			nop->Mf = op->Mf;
			nop->ncoil = op->ncoil;
			if (nop->K > 0) {
				try {
					nop->bksp.alloc((size_t)op->bksp.n);
				} catch (...) {
					delete nop;
					delete op;
					throw;
				}
			}
			*fft_opp = (const struct linop_s*)nop;
		}

		return (const struct linop_s*)op;
	}, nullptr);
}

void linop_free(const struct linop_s* _op)
{
	guarded_void([&]() {
		Op* op = (Op*)_op;
		if (nullptr != op) {
			BAM_HIP(hipStreamSynchronize(op->stream));
			delete op;
		}
	});
}

void linop_forward(const struct linop_s* _op, int DN, const long ddims[], float _Complex* dst, int SN,
		   const long sdims[], const float _Complex* src)
{
	guarded_void([&]() {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "linop_forward: NULL operator");
		check_dims("linop_forward (output)", DN, ddims, op->N, op->odims);
		check_dims("linop_forward (input)", SN, sdims, op->N, op->idims);

		apply_staged(op, (float2*)dst, op->osize, (const float2*)src, op->isize,
			     [&](float2* d, const float2* s) { op->fwd(d, s); });
	});
}

void linop_adjoint(const struct linop_s* _op, int DN, const long ddims[], float _Complex* dst, int SN,
		   const long sdims[], const float _Complex* src)
{
	guarded_void([&]() {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "linop_adjoint: NULL operator");
		check_dims("linop_adjoint (output)", DN, ddims, op->N, op->idims);
		check_dims("linop_adjoint (input)", SN, sdims, op->N, op->odims);

		apply_staged(op, (float2*)dst, op->isize, (const float2*)src, op->osize,
			     [&](float2* d, const float2* s) { op->adj(d, s); });
	});
}

void linop_normal(const struct linop_s* _op, int N, const long dims[], float _Complex* dst, const float _Complex* src)
{
	guarded_void([&]() {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "linop_normal: NULL operator");
		check_dims("linop_normal", N, dims, op->N, op->idims);

		op->prepare_normal();
		apply_staged(op, (float2*)dst, op->isize, (const float2*)src, op->isize,
			     [&](float2* d, const float2* s) { op->nrm(d, s, nullptr); });
	});
}

int nufft_get_psf_dims(const struct linop_s* _op, int N, long psf_dims[])
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "nufft_get_psf_dims: NULL operator");
		op->plan("nufft_get_psf_dims");

		if (N > 0) {
			BAM_CHECK(N == op->N + 1, "nufft_get_psf_dims: need %d dims", op->N + 1);
			psf_dims_of(op, psf_dims);
		}

		return op->N + 1;
	}, -1);
}

void nufft_get_psf2(const struct linop_s* _op, int N, const long psf_dims[], const long psf_strs[],
		    float _Complex* psf)
{
	guarded_void([&]() {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op && op->plan("nufft_get_psf").has_psf, "nufft_get_psf: operator has no PSF");
		long want[33];
		BAM_CHECK(N == op->N + 1, "nufft_get_psf: need %d dims", op->N + 1);
		psf_dims_of(op, want);
		check_dims("nufft_get_psf", N, psf_dims, N, want);
		std::lock_guard<std::mutex> lock(op->mtx());
		copy_strides(N, psf_dims, psf_strs, (float2*)psf, op->nufft->psf.ptr, true, op->stream);
	});
}

void nufft_get_psf(const struct linop_s* _op, int N, const long psf_dims[], float _Complex* psf)
{
	long strs[33];
	long d = sizeof(float2);

	for (int i = 0; i < N && i < 33; i++) {
		strs[i] = d;
		d *= psf_dims[i];
	}

	nufft_get_psf2(_op, N, psf_dims, strs, psf);
}

void nufft_update_psf2(const struct linop_s* _op, int ND, const long psf_dims[], const long psf_strs[],
		       const float _Complex* psf)
{
	guarded_void([&]() {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "nufft_update_psf: NULL operator");
		BAM_CHECK(op->plan("nufft_update_psf").toeplitz, "nufft_update_psf: operator is not in Toeplitz mode");
		long want[33];
		BAM_CHECK(ND == op->N + 1, "nufft_update_psf: dims");
		psf_dims_of(op, want);
		check_dims("nufft_update_psf", ND, psf_dims, ND, want);
		std::lock_guard<std::mutex> lock(op->mtx());
		long n = (long)op->nufft->S * op->nufft->N;
		if ((long)op->nufft->psf.n < n)
			op->nufft->psf.alloc(n);
		copy_strides(ND, psf_dims, psf_strs, (float2*)psf, op->nufft->psf.ptr, false, op->stream);
		op->nufft->apply_real_psf();
		op->nufft->has_psf = true;
		if (op->cg)
			op->cg->drop_graph();
	});
}

void nufft_update_psf(const struct linop_s* _op, int ND, const long psf_dims[], const float _Complex* psf)
{
	long strs[33];
	long d = sizeof(float2);

	for (int i = 0; i < ND && i < 33; i++) {
		strs[i] = d;
		d *= psf_dims[i];
	}

	nufft_update_psf2(_op, ND, psf_dims, strs, psf);
}

struct linop_s* nufft_create_normal(int N, const long cim_dims[], int ND, const long psf_dims[],
				    const float _Complex* psf, bool basis, struct nufft_conf_s conf)
{
	return guarded([&]() -> struct linop_s* {
		check_conf(conf, basis);
		BAM_CHECK(conf.toeplitz, "nufft_create_normal: needs the Toeplitz mode");
		BAM_CHECK(N >= 4 && N <= 32 && ND == N + 1, "nufft_create_normal: dims (N = %d, ND = %d)", N, ND);

		for (int i = 4; i < N; i++)
			BAM_CHECK(1 == cim_dims[i], "nufft_create_normal: dim %d must be 1", i);

		int img[3];

		for (int i = 0; i < 3; i++) {
			BAM_CHECK(cim_dims[i] >= 1 && cim_dims[i] < (1l << 20), "nufft_create_normal: bad image dim");
			img[i] = (int)cim_dims[i];
		}

		auto op = new Op();
		op->kind = OP_NUFFT;
		op->N = N;
		op->normal_only = true;

		// BART's normal-only linop has singleton output dims
		for (int i = 0; i < N; i++)
			op->odims[i] = 1;
		memcpy(op->idims, cim_dims, sizeof(long) * N);
		op->osize = 1;
		op->isize = md_size(N, cim_dims);
		op->q = std::make_shared<Queue>();
		op->stream = op->q->stream;

		try {
			op->nufft = std::make_shared<Nufft>();
			op->nufft->stream = op->stream;
			op->nufft->create_geometry(img, (int)cim_dims[3], true, conf.periodic, conf.width, conf.os);
			long want[33];
			psf_dims_of(op, want);
			check_dims("nufft_create_normal (psf)", ND, psf_dims, ND, want);
			op->nufft->psf.alloc((size_t)op->nufft->S * op->nufft->N);
		} catch (...) {
			delete op;
			throw;
		}

		if (nullptr != psf) {
			long strs[33];
			long d = sizeof(float2);
			for (int i = 0; i < ND; i++) {
				strs[i] = d;
				d *= psf_dims[i];
			}
			try {
				copy_strides(ND, psf_dims, strs, (float2*)psf, op->nufft->psf.ptr, false, op->stream);
			} catch (...) {
				delete op;
				throw;
			}
			op->nufft->has_psf = true;
		}

		return (struct linop_s*)op;
	}, nullptr);
}

void nufft_update_traj(const struct linop_s* _op, int N, const long trj_dims[], const float _Complex* traj,
		       const long wgh_dims[], const float _Complex* weights, const long bas_dims[],
		       const float _Complex* basis)
{
	(void)bas_dims;
	guarded_void([&]() {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "nufft_update_traj: NULL operator");
		Nufft& p = op->plan("nufft_update_traj");
		BAM_CHECK(!op->normal_only, "nufft_update_traj: normal-only operator has no trajectory");
		BAM_CHECK(nullptr == basis, "nufft_update_traj: subspace basis is not supported by this library");
		BAM_CHECK(nullptr != traj, "nufft_update_traj: trajectory is NULL");
		BAM_CHECK(N == op->N, "nufft_update_traj: dimension count %d != %d", N, op->N);
		BAM_CHECK(3 == trj_dims[0] && trj_dims[1] == op->odims[1] && trj_dims[2] == op->odims[2],
			  "Incompatible dimensions of k-space and trajectory!");

		for (int i = 3; i < N; i++)
			BAM_CHECK(1 == trj_dims[i], "nufft_update_traj: trajectory dim %d must be 1", i);

		const long M = trj_dims[1] * trj_dims[2];
		std::vector<float> tr = traj_reals((const float2*)traj, M);
		std::vector<float2> w = expand_weights(N, wgh_dims, (const float2*)weights, op->odims);

		std::lock_guard<std::mutex> lock(op->mtx());
		BAM_HIP(hipStreamSynchronize(op->stream));
		p.set_traj(M, tr.data(), w.empty() ? nullptr : w.data(), p.toeplitz);
		if (op->cg)
			op->cg->drop_graph();
	});
}

float _Complex* compute_psf2(int N, const long psf_dims[], unsigned long flags, const long trj_dims[],
			     const float _Complex* traj, const long bas_dims[], const float _Complex* basis,
			     const long wgh_dims[], const float _Complex* weights, bool periodic, bool lowmem,
			     bool upper_triag)
{
	(void)bas_dims;
	(void)lowmem;
	return guarded([&]() -> float _Complex* {
		(void)upper_triag;  // no basis: the PSF does not depend on it
		BAM_CHECK(nullptr == basis, "compute_psf2: subspace basis is not supported by this library");
		BAM_CHECK(7 == (flags & 7) && 0 == (flags & ~7ul), "compute_psf2: flags must be 7");
		BAM_CHECK(N >= 4 && N <= 32, "compute_psf2: unsupported dimension count %d", N);
		BAM_CHECK(nullptr != traj, "compute_psf2: trajectory is NULL");
		BAM_CHECK(3 == trj_dims[0], "compute_psf2: trajectory dim 0 must be 3");

		for (int i = 3; i <= N; i++)
			BAM_CHECK(1 == trj_dims[i] && (i == N || 1 == psf_dims[i]), "compute_psf2: dim %d must be 1", i);

		int img[3];
		for (int i = 0; i < 3; i++)
			img[i] = (int)psf_dims[i];

		const long M = trj_dims[1] * trj_dims[2];
		std::vector<float> tr = traj_reals((const float2*)traj, M);
		long ksp_dims[33];
		for (int i = 0; i <= N; i++)
			ksp_dims[i] = (0 == i) ? 1 : trj_dims[i];
		std::vector<float2> w = expand_weights(N, wgh_dims, (const float2*)weights, ksp_dims);

		Queue q;
		Nufft p;
		p.stream = q.stream;
		p.create_geometry(img, 1, true, periodic, nufft_conf_defaults.width, nufft_conf_defaults.os);
		BAM_CHECK(psf_dims[N] == p.S, "compute_psf2: psf dim %d must be %d", N, p.S);
		p.M = M;
		p.compute_psf_from_traj(tr.data(), w.empty() ? nullptr : w.data(), periodic);

		// the result lives where the trajectory does (BART: md_alloc_sameplace)
		long n = (long)p.S * p.N;
		float2* out = nullptr;
		if (on_device(traj)) {
			BAM_HIP(hipMalloc((void**)&out, sizeof(float2) * n));
			BAM_HIP(hipMemcpy(out, p.psf.ptr, sizeof(float2) * n, hipMemcpyDeviceToDevice));
		} else {
			out = (float2*)malloc(sizeof(float2) * n);
			BAM_CHECK(nullptr != out, "compute_psf2: out of memory");
			BAM_HIP(hipMemcpy(out, p.psf.ptr, sizeof(float2) * n, hipMemcpyDeviceToHost));
		}
		return (float _Complex*)out;
	}, nullptr);
}

int bart_amd_linop_set_stream(const struct linop_s* _op, void* stream)
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "bart_amd_linop_set_stream: NULL operator");
		std::lock_guard<std::mutex> lock(op->mtx());
		op->user_stream = (hipStream_t)stream;
		return 0;
	}, -1);
}

static void fft_common(int D, const long dims[], unsigned long flags, float _Complex* dst, const float _Complex* src,
		       float sign)
{
	guarded_void([&]() {
		BAM_CHECK(D >= 1, "fft: bad dims");
		BAM_CHECK(0 == (flags & ~7ul), "fft: only dims 0-2 can be transformed");
		int d3[3] = { 1, 1, 1 };
		for (int i = 0; i < D && i < 3; i++)
			d3[i] = (int)dims[i];
		long batch = 1;
		for (int i = 3; i < D; i++)
			batch *= dims[i];
		long n = (long)d3[0] * d3[1] * d3[2] * batch;

		for (int a = 0; a < 3; a++)
			if ((flags & (1ul << a)) && d3[a] > 1)
				BAM_CHECK(fft_size_supported(d3[a]), "fft: length %d not supported", d3[a]);

		bool dd = on_device(dst), sd = on_device(src);
		hipStream_t st = 0;

		if (dd && sd) {
			fft_nd(d3, batch, (unsigned)flags, sign, (float2*)dst, (const float2*)src, 1.f, st);
			BAM_HIP(hipStreamSynchronize(st));
			return;
		}

		DevBuf<float2> buf(n);
		BAM_HIP(hipMemcpy(buf.ptr, src, sizeof(float2) * n, sd ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
		fft_nd(d3, batch, (unsigned)flags, sign, buf.ptr, buf.ptr, 1.f, st);
		BAM_HIP(hipMemcpy(dst, buf.ptr, sizeof(float2) * n, dd ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
	});
}

void fft(int D, const long dimensions[], unsigned long flags, float _Complex* dst, const float _Complex* src)
{
	fft_common(D, dimensions, flags, dst, src, -1.f);
}

void ifft(int D, const long dimensions[], unsigned long flags, float _Complex* dst, const float _Complex* src)
{
	fft_common(D, dimensions, flags, dst, src, +1.f);
}

int lsqr_conjgrad(const struct linop_s* _op, int maxiter, float l2lambda, float tol, int N, const long img_dims[],
		  float _Complex* x, const float _Complex* b, bool warmstart, bool warm_fallback, float* resid_trace,
		  struct cg_stats_s* stats)
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "lsqr_conjgrad: NULL operator");
		check_dims("lsqr_conjgrad", N, img_dims, op->N, op->idims);
		BAM_CHECK(!op->toeplitz() || op->nufft->has_psf, "lsqr_conjgrad: Toeplitz operator without PSF");

		std::lock_guard<std::mutex> lock(op->mtx());

		// x and b may have been written on the caller's stream
		if (0 != op->user_stream)
			BAM_HIP(hipStreamSynchronize(op->user_stream));

		long ne = op->isize;

		if (!op->cg)
			op->cg.reset(new CgSolver(2 * ne, op->stream));

		// the k-space buffer of an explicit normal operator exists before
		// the iterations are captured
		op->prepare_normal();
		NormalFn normal = [op](float2* d, const float2* s, const int* skip, const CgApFuse* f) {
			return op->nrm(d, s, skip, f);
		};

		bool xd = on_device(x), bd = on_device(b);
		float2* xdev = (float2*)x;
		const float2* bdev = (const float2*)b;
		DevBuf<float2> xs, bs;

		if (!xd) {
			xs.alloc(ne);
			BAM_HIP(hipMemcpy(xs.ptr, x, sizeof(float2) * ne, hipMemcpyHostToDevice));
			xdev = xs.ptr;
		}

		if (!bd) {
			bs.alloc(ne);
			BAM_HIP(hipMemcpy(bs.ptr, b, sizeof(float2) * ne, hipMemcpyHostToDevice));
			bdev = bs.ptr;
		}

		CgResult r = op->cg->solve(normal, l2lambda, maxiter, tol, xdev, bdev, warmstart, warm_fallback, true);

		if (!xd)
			BAM_HIP(hipMemcpy(x, xdev, sizeof(float2) * ne, hipMemcpyDeviceToHost));

		if (nullptr != resid_trace && r.iterations > 0)
			memcpy(resid_trace, r.trace.data(), sizeof(float) * r.iterations);

		if (nullptr != op->mon_err)
			op->cg->bt.monitor_errors(op->mon_err, r.iterations);

		if (nullptr != stats) {
			stats->iterations = r.iterations;
			stats->status = r.status;
			stats->resid = r.resid;
			stats->bnorm = r.bnorm;
			stats->time_ms = r.ms;
			stats->fell_back = r.fell_back;
		}

		return r.iterations;
	}, -1);
}

int bart_amd_cg_monitor(const struct linop_s* _op, const float _Complex* truth, float* err)
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "bart_amd_cg_monitor: NULL operator");
		std::lock_guard<std::mutex> lock(op->mtx());
		const long ne = op->isize;

		if (!op->cg)
			op->cg.reset(new CgSolver(2 * ne, op->stream));

		if (nullptr == truth) {
			op->cg->set_monitor(nullptr, 0.);
			op->mon_err = nullptr;
			return 0;
		}

		BAM_CHECK(nullptr != err, "bart_amd_cg_monitor: NULL error array");
		std::vector<float2> h = to_host((const float2*)truth, ne);
		double s = 0.;
		for (const float2& v : h)
			s += (double)v.x * v.x + (double)v.y * v.y;
		DevBuf<float2> d(ne);
		BAM_HIP(hipMemcpy(d.ptr, h.data(), sizeof(float2) * ne, hipMemcpyHostToDevice));
		op->cg->set_monitor(d.ptr, sqrt(s));
		op->mon_err = err;
		return 0;
	}, -1);
}

int lsqr_conjgrad_noalloc(const struct linop_s* _op, int maxiter, float l2lambda, float tol, int N,
			  const long img_dims[], float _Complex* x, const float _Complex* b, bool warmstart,
			  bool warm_fallback, float* resid_trace, struct cg_stats_s* stats)
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op, "lsqr_conjgrad_noalloc: NULL operator");
		check_dims("lsqr_conjgrad_noalloc", N, img_dims, op->N, op->idims);
		BAM_CHECK(!op->toeplitz() || op->nufft->has_psf, "lsqr_conjgrad_noalloc: Toeplitz operator without PSF");
		BAM_CHECK(op->toeplitz() || (long)op->kbuf.n >= op->osize,
			  "lsqr_conjgrad_noalloc: the normal operator adjoint(forward(x)) needs its k-space buffer "
			  "(create the operator with conf.toeplitz, or call lsqr_conjgrad on it once)");
		BAM_CHECK(on_device(x) && on_device(b), "lsqr_conjgrad_noalloc: x and b must be device pointers");

		std::lock_guard<std::mutex> lock(op->mtx());

		// x and b may have been written on the caller's stream
		if (0 != op->user_stream)
			BAM_HIP(hipStreamSynchronize(op->user_stream));

		NormalFn normal = [op](float2* d, const float2* s, const int* skip, const CgApFuse* f) {
			return op->nrm(d, s, skip, f);
		};
		CgResult r = cg_solve_noalloc(normal, l2lambda, maxiter, tol, (float2*)x, (const float2*)b, 2 * op->isize,
					      warmstart, warm_fallback, op->stream, resid_trace);

		if (nullptr != stats) {
			stats->iterations = r.iterations;
			stats->status = r.status;
			stats->resid = r.resid;
			stats->bnorm = r.bnorm;
			stats->time_ms = 0.f;
			stats->fell_back = r.fell_back;
		}

		return r.iterations;
	}, -1);
}

// Streams the problems of a batched solve are spread over; the device has
// four hardware queues per process by default.
constexpr int CG_BATCH_STREAMS = 4;

int lsqr_conjgrad_batch(int nprob, const struct linop_s* const ops[], int maxiter, float l2lambda, float tol, int N,
			float _Complex* const x[], const float _Complex* const b[], bool warmstart, bool warm_fallback,
			float* resid_trace, struct cg_stats_s stats[])
{
	struct Workers {

		std::vector<hipStream_t> ws;
		std::vector<hipEvent_t> ev;
		std::vector<Op*> ops;

		~Workers()
		{
			for (hipStream_t s : ws)
				(void)hipStreamSynchronize(s);
			for (Op* o : ops)
				o->set_plan_stream(o->stream);
			for (hipEvent_t e : ev)
				(void)hipEventDestroy(e);
			for (hipStream_t s : ws)
				(void)hipStreamDestroy(s);
		}
	};

	return guarded([&]() -> int {
		BAM_CHECK(nprob >= 0, "lsqr_conjgrad_batch: negative problem count");

		if (0 == nprob)
			return 0;

		BAM_CHECK(nullptr != ops && nullptr != x && nullptr != b, "lsqr_conjgrad_batch: NULL argument");

		std::vector<Op*> op(nprob);
		std::vector<const void*> plans(nprob);
		std::vector<std::mutex*> locks;

		for (int i = 0; i < nprob; i++) {

			op[i] = (Op*)ops[i];
			BAM_CHECK(nullptr != op[i], "lsqr_conjgrad_batch: NULL operator (problem %d)", i);
			BAM_CHECK(N == op[i]->N, "lsqr_conjgrad_batch: dimension count %d != %d (problem %d)", N, op[i]->N, i);
			BAM_CHECK(!op[i]->toeplitz() || op[i]->nufft->has_psf,
				  "lsqr_conjgrad_batch: Toeplitz operator without PSF (problem %d)", i);
			BAM_CHECK(nullptr != x[i] && nullptr != b[i], "lsqr_conjgrad_batch: NULL vector (problem %d)", i);
			plans[i] = op[i]->plan_id();
			locks.push_back(&op[i]->mtx());
		}

		// the problems' normal operators run concurrently: no two may share a plan
		std::vector<const void*> ps = plans;
		std::sort(ps.begin(), ps.end());
		BAM_CHECK(std::adjacent_find(ps.begin(), ps.end()) == ps.end(),
			  "lsqr_conjgrad_batch: two problems share one NUFFT plan");

		std::sort(locks.begin(), locks.end());
		locks.erase(std::unique(locks.begin(), locks.end()), locks.end());
		std::vector<std::unique_lock<std::mutex>> held;
		for (std::mutex* m : locks)
			held.emplace_back(*m);

		// device copies of host vectors
		std::vector<std::unique_ptr<DevBuf<float2>>> stage;
		std::vector<float2*> xd(nprob);
		std::vector<const float2*> bd(nprob);
		std::vector<char> xhost(nprob, 0);

		for (int i = 0; i < nprob; i++) {

			long ne = op[i]->isize;
			xd[i] = (float2*)x[i];
			bd[i] = (const float2*)b[i];

			if (!on_device(x[i])) {
				stage.emplace_back(new DevBuf<float2>(ne));
				BAM_HIP(hipMemcpy(stage.back()->ptr, x[i], sizeof(float2) * ne, hipMemcpyHostToDevice));
				xd[i] = stage.back()->ptr;
				xhost[i] = 1;
			}

			if (!on_device(b[i])) {
				stage.emplace_back(new DevBuf<float2>(ne));
				BAM_HIP(hipMemcpy(stage.back()->ptr, b[i], sizeof(float2) * ne, hipMemcpyHostToDevice));
				bd[i] = stage.back()->ptr;
			}
		}

		// When every problem is a 2-D Toeplitz SENSE operator, one stream
		// whose iteration runs the normal operators of all problems grouped
		// by image size, one launch per pass and size
		// (BART_AMD_BATCH_GROUPS=off: the problems spread over streams instead)
		static const bool groups_on = [] {
			const char* e = getenv("BART_AMD_BATCH_GROUPS");
			return !(nullptr != e && 0 == strcmp(e, "off"));
		}();
		bool grouped = groups_on && nprob > 1;

		for (int i = 0; i < nprob && grouped; i++)
			grouped = OP_SENSE == op[i]->kind && 1 == op[i]->nchunks() && nullptr != op[i]->nufft &&
				  1 == op[i]->nbatch && normal_groupable(*op[i]->nufft, op[i]->maps());

		// longest-processing-time assignment of the problems to the streams
		const int K = grouped ? 1 : std::min(nprob, CG_BATCH_STREAMS);
		std::vector<int> order(nprob);
		for (int i = 0; i < nprob; i++)
			order[i] = i;
		std::stable_sort(order.begin(), order.end(),
				 [&](int a, int c) { return op[a]->cost() > op[c]->cost(); });

		std::vector<std::vector<int>> members(K);
		std::vector<double> load(K, 0.);

		for (int i : order) {
			int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
			members[k].push_back(i);
			load[k] += op[i]->cost();
		}

		// (declared first: destroyed after the streams have drained)
		std::vector<std::unique_ptr<CgBatch>> bt(K);
		Workers w;

		for (int k = 0; k < K; k++) {

			hipStream_t s;
			BAM_HIP(hipStreamCreate(&s));
			w.ws.push_back(s);
			hipEvent_t e0, e1;
			BAM_HIP(hipEventCreate(&e0));
			w.ev.push_back(e0);
			BAM_HIP(hipEventCreate(&e1));
			w.ev.push_back(e1);

			std::vector<long> nf;

			for (int i : members[k]) {
				op[i]->prepare_normal();
				BAM_HIP(hipStreamSynchronize(op[i]->stream));
				op[i]->set_plan_stream(s);
				w.ops.push_back(op[i]);
				nf.push_back(2 * op[i]->isize);
			}

			bt[k].reset(new CgBatch(nf, s));
		}

		// the grouped normal operators of a single batch (grouped)
		std::shared_ptr<NormalGroupSet> gset;
		CgGroupHooks hooks;

		if (grouped) {

			const std::vector<int>& mem = members[0];

			hooks.prepare = [&op, &mem, &gset](CgBatch& b, std::vector<CgGrpItem>& items) {
				std::vector<const Nufft*> plans;
				std::vector<const float2*> sens, src;
				std::vector<float2*> dst;
				std::vector<const int*> skip;

				for (size_t j = 0; j < mem.size(); j++) {
					const Op* o = op[mem[j]];
					plans.push_back(o->nufft.get());
					sens.push_back(o->maps());
					src.push_back((const float2*)(b.p.ptr + b.seg_h[j].off));
					dst.push_back((float2*)(b.Ap.ptr + b.seg_h[j].off));
					skip.push_back(&b.st.ptr[j].done);
				}

				std::vector<int> ng;
				gset = normal_groups_build(plans, sens, src, dst, skip, ng);

				for (size_t j = 0; j < mem.size(); j++) {
					auto pr = normal_group_partials(*plans[j], dst[j]);
					items[j] = CgGrpItem{ (const float*)pr.first, 2 * plans[j]->N, pr.second };
				}
			};

			hooks.launch = [&gset](CgBatch& b) { normal_groups_launch(*gset, b.stream); };
			bt[0]->hooks = &hooks;
		}

		std::vector<BatchNormalFn> fn(K);

		for (int k = 0; k < K; k++) {

			const std::vector<int>& mem = members[k];
			fn[k] = [&op, &mem](int j, float2* d, const float2* s, const int* skip, const CgApFuse* f) {
				return op[mem[j]]->nrm(d, s, skip, f);
			};

			std::vector<const float2*> x0, bb;
			for (int i : mem) {
				x0.push_back(xd[i]);
				bb.push_back(bd[i]);
			}

			BAM_HIP(hipEventRecord(w.ev[2 * k], w.ws[k]));
			bt[k]->start(fn[k], l2lambda, tol, maxiter, x0.data(), bb.data(), warmstart, warm_fallback);
		}

		{
			std::vector<CgBatch*> bp;
			std::vector<const BatchNormalFn*> fp;
			for (int k = 0; k < K; k++) {
				bp.push_back(bt[k].get());
				fp.push_back(&fn[k]);
			}
			iterate_batches(bp, fp, l2lambda, maxiter);
		}

		for (int k = 0; k < K; k++) {

			std::vector<float2*> xo;
			for (int i : members[k])
				xo.push_back(xd[i]);

			bt[k]->finish(xo.data());
			BAM_HIP(hipEventRecord(w.ev[2 * k + 1], w.ws[k]));
		}

		int maxit = 0;

		for (int k = 0; k < K; k++) {

			BAM_HIP(hipStreamSynchronize(w.ws[k]));
			float ms = 0.f;
			BAM_HIP(hipEventElapsedTime(&ms, w.ev[2 * k], w.ev[2 * k + 1]));

			for (size_t j = 0; j < members[k].size(); j++) {

				const int i = members[k][j];
				CgResult r = bt[k]->result((int)j);
				maxit = std::max(maxit, r.iterations);

				if (xhost[i])
					BAM_HIP(hipMemcpy(x[i], xd[i], sizeof(float2) * op[i]->isize, hipMemcpyDeviceToHost));

				if (nullptr != resid_trace && r.iterations > 0)
					memcpy(resid_trace + (long)i * std::max(1, maxiter), r.trace.data(),
					       sizeof(float) * r.iterations);

				if (nullptr != stats) {
					stats[i].iterations = r.iterations;
					stats[i].status = r.status;
					stats[i].resid = r.resid;
					stats[i].bnorm = r.bnorm;
					stats[i].time_ms = ms;
					stats[i].fell_back = r.fell_back;
				}
			}
		}

		return maxit;
	}, -1);
}

// ---------------------------------------------------------------------------
// Layers of the CG warm-start network (no counterpart in BART)
// ---------------------------------------------------------------------------

static void check_dev(const char* what, const void* p)
{
	BAM_CHECK(nullptr == p || on_device(p), "%s: pointers must be device memory", what);
}

int bart_amd_conv3x3(const float* x, int B, int Cin, int H, int W, const float* w, const float* bias,
		     const float* res, int Cout, int relu, float* y)
{
	return guarded([&]() -> int {
		BAM_CHECK(nullptr != x && nullptr != w && nullptr != y, "bart_amd_conv3x3: NULL argument");
		for (const void* p : { (const void*)x, (const void*)w, (const void*)bias, (const void*)res, (const void*)y })
			check_dev("bart_amd_conv3x3", p);
		conv3x3_fwd(x, B, Cin, H, W, w, bias, res, Cout, 0 != relu, y, default_stream());
		return 0;
	}, -1);
}

int bart_amd_conv3x3_bwd_data(const float* dy, int B, int Cout, int H, int W, const float* w, int Cin, float* dx)
{
	return guarded([&]() -> int {
		BAM_CHECK(nullptr != dy && nullptr != w && nullptr != dx, "bart_amd_conv3x3_bwd_data: NULL argument");
		for (const void* p : { (const void*)dy, (const void*)w, (const void*)dx })
			check_dev("bart_amd_conv3x3_bwd_data", p);
		conv3x3_bwd_data(dy, B, Cout, H, W, w, Cin, dx, default_stream());
		return 0;
	}, -1);
}

int bart_amd_conv3x3_bwd_weights(const float* x, const float* dy, int B, int Cin, int Cout, int H, int W, float* dw,
				 float* db)
{
	return guarded([&]() -> int {
		BAM_CHECK(nullptr != x && nullptr != dy && nullptr != dw, "bart_amd_conv3x3_bwd_weights: NULL argument");
		for (const void* p : { (const void*)x, (const void*)dy, (const void*)dw, (const void*)db })
			check_dev("bart_amd_conv3x3_bwd_weights", p);
		conv3x3_bwd_weights(x, dy, B, Cin, Cout, H, W, dw, db, default_stream());
		return 0;
	}, -1);
}

struct bart_amd_warmnet* bart_amd_warmnet_create(int channels, unsigned long seed)
{
	return guarded([&]() -> struct bart_amd_warmnet* {
		return (struct bart_amd_warmnet*)new WarmNet(channels, seed);
	}, nullptr);
}

void bart_amd_warmnet_free(struct bart_amd_warmnet* net)
{
	guarded_void([&]() {
		BAM_HIP(hipStreamSynchronize(default_stream()));
		delete (WarmNet*)net;
	});
}

long bart_amd_warmnet_num_params(const struct bart_amd_warmnet* net)
{
	return guarded([&]() -> long {
		BAM_CHECK(nullptr != net, "bart_amd_warmnet_num_params: NULL network");
		return ((const WarmNet*)net)->nparams;
	}, -1l);
}

int bart_amd_warmnet_params(struct bart_amd_warmnet* _net, float* host, int set)
{
	return guarded([&]() -> int {
		WarmNet* net = (WarmNet*)_net;
		BAM_CHECK(nullptr != net && nullptr != host, "bart_amd_warmnet_params: NULL argument");
		BAM_CHECK(set >= 0 && set <= 2, "bart_amd_warmnet_params: mode 0, 1 or 2");
		BAM_HIP(hipStreamSynchronize(default_stream()));
		if (1 == set)
			BAM_HIP(hipMemcpy(net->params.ptr, host, sizeof(float) * net->nparams, hipMemcpyHostToDevice));
		else
			BAM_HIP(hipMemcpy(host, (0 == set) ? net->params.ptr : net->grads.ptr, sizeof(float) * net->nparams,
					  hipMemcpyDeviceToHost));
		return 0;
	}, -1);
}

int bart_amd_warmnet_forward(struct bart_amd_warmnet* net, const float* x, int B, int H, int W, float* y)
{
	return guarded([&]() -> int {
		BAM_CHECK(nullptr != net && nullptr != x && nullptr != y, "bart_amd_warmnet_forward: NULL argument");
		check_dev("bart_amd_warmnet_forward", x);
		check_dev("bart_amd_warmnet_forward", y);
		((WarmNet*)net)->forward(x, B, H, W, y, default_stream());
		return 0;
	}, -1);
}

double bart_amd_warmnet_train_step(struct bart_amd_warmnet* net, const float* x, const float* target, int B, int H,
				   int W, float lr)
{
	return guarded([&]() -> double {
		BAM_CHECK(nullptr != net && nullptr != x && nullptr != target, "bart_amd_warmnet_train_step: NULL argument");
		check_dev("bart_amd_warmnet_train_step", x);
		check_dev("bart_amd_warmnet_train_step", target);
		return ((WarmNet*)net)->train_step(x, target, B, H, W, lr, default_stream());
	}, -1.);
}

int bart_amd_warmnet_train_step_grad(struct bart_amd_warmnet* net, const float* x, const float* dout, int B, int H,
				     int W, float lr)
{
	return guarded([&]() -> int {
		BAM_CHECK(nullptr != net && nullptr != x && nullptr != dout,
			  "bart_amd_warmnet_train_step_grad: NULL argument");
		check_dev("bart_amd_warmnet_train_step_grad", x);
		check_dev("bart_amd_warmnet_train_step_grad", dout);
		((WarmNet*)net)->train_step_grad(x, dout, B, H, W, lr, default_stream());
		return 0;
	}, -1);
}

// the step along d (d = b when NULL: CG's first iterate from zero), its
// scale and the network channels of x1 / s in op->g_in (b, d: host or
// device)
static void warm_input(Op* op, float l2lambda, int N, const long img_dims[], const float _Complex* b,
		       const float _Complex* d, const char* what)
{
	check_dims(what, N, img_dims, op->N, op->idims);
	BAM_CHECK(OP_NUDFT != op->kind && nullptr != op->nufft, "%s: not for the chained NUFFT", what);
	BAM_CHECK(1 == op->nufft->n[2] && op->nufft->n[0] > 1 && op->nufft->n[1] > 1,
		  "%s: 2-D images only", what);
	BAM_CHECK(0 == op->K, "%s: not for an operator with a subspace basis", what);
	BAM_CHECK(1 == op->nbatch, "%s: not for an operator with batch dims", what);
	BAM_CHECK(!op->toeplitz() || op->nufft->has_psf, "%s: Toeplitz operator without PSF", what);
	const long npix = op->isize;

	if ((long)op->g_ab.n < npix) {
		op->g_ab.alloc(npix);
		op->g_in.alloc(2 * npix);
		op->g_out.alloc(2 * npix);
	}
	if (op->g_part.n < (size_t)GUESS_PARTS) {
		op->g_part.alloc(GUESS_PARTS);
		op->g_sc.alloc(4);
	}

	const float2* bd = (const float2*)b;

	if (!on_device(b)) {
		if ((long)op->g_b.n < npix)
			op->g_b.alloc(npix);
		BAM_HIP(hipMemcpyAsync(op->g_b.ptr, b, sizeof(float2) * npix, hipMemcpyHostToDevice, op->stream));
		bd = op->g_b.ptr;
	}

	const float2* dd = bd;

	if (nullptr != d) {
		dd = (const float2*)d;
		if (!on_device(d)) {
			if ((long)op->g_d.n < npix)
				op->g_d.alloc(npix);
			BAM_HIP(hipMemcpyAsync(op->g_d.ptr, d, sizeof(float2) * npix, hipMemcpyHostToDevice, op->stream));
			dd = op->g_d.ptr;
		}
	}

	op->prepare_normal();
	op->nrm(op->g_ab.ptr, dd, nullptr);
	guess_scalars(dd, bd, op->g_ab.ptr, l2lambda, npix, op->g_part.ptr, op->g_sc.ptr, op->stream);
	guess_pack(dd, op->g_sc.ptr, 1.f, npix, true, op->g_in.ptr, op->stream);
}

int bart_amd_warmnet_input_dir(const struct linop_s* _op, float l2lambda, int N, const long img_dims[],
			       const float _Complex* b, const float _Complex* d, float* inp, float* scale)
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op && nullptr != b && nullptr != inp, "bart_amd_warmnet_input: NULL argument");
		check_dev("bart_amd_warmnet_input", inp);
		std::lock_guard<std::mutex> lock(op->mtx());
		if (0 != op->user_stream)
			BAM_HIP(hipStreamSynchronize(op->user_stream));
		warm_input(op, l2lambda, N, img_dims, b, d, "bart_amd_warmnet_input");
		BAM_HIP(hipMemcpyAsync(inp, op->g_in.ptr, sizeof(float) * 2 * op->isize, hipMemcpyDeviceToDevice,
				       op->stream));
		float sc[3];
		BAM_HIP(hipMemcpyAsync(sc, op->g_sc.ptr, sizeof(float) * 3, hipMemcpyDeviceToHost, op->stream));
		BAM_HIP(hipStreamSynchronize(op->stream));
		if (nullptr != scale)
			*scale = sc[1];
		return 0;
	}, -1);
}

int bart_amd_warmnet_input(const struct linop_s* op, float l2lambda, int N, const long img_dims[],
			   const float _Complex* b, float* inp, float* scale)
{
	return bart_amd_warmnet_input_dir(op, l2lambda, N, img_dims, b, nullptr, inp, scale);
}

int bart_amd_warmnet_pack(int N, const long img_dims[], const float _Complex* img, float scale, float* out)
{
	return guarded([&]() -> int {
		BAM_CHECK(nullptr != img && nullptr != out && scale > 0.f, "bart_amd_warmnet_pack: bad argument");
		check_dev("bart_amd_warmnet_pack", img);
		check_dev("bart_amd_warmnet_pack", out);
		const long npix = md_size(N, img_dims);
		guess_pack((const float2*)img, nullptr, 1.f / scale, npix, false, out, default_stream());
		BAM_HIP(hipStreamSynchronize(default_stream()));
		return 0;
	}, -1);
}

// the guess along d (NULL: b) for an operator whose lock the caller holds
static void warm_guess(WarmNet* net, Op* op, float l2lambda, int N, const long img_dims[], const float _Complex* b,
		       const float _Complex* d, float _Complex* x0)
{
	warm_input(op, l2lambda, N, img_dims, b, d, "bart_amd_warmnet_guess");
	net->forward(op->g_in.ptr, 1, op->nufft->n[1], op->nufft->n[0], op->g_out.ptr, op->stream);

	const long npix = op->isize;
	float2* xd = (float2*)x0;
	const bool xdev = on_device(x0);

	if (!xdev) {
		if ((long)op->g_x.n < npix)
			op->g_x.alloc(npix);
		xd = op->g_x.ptr;
	}

	guess_unpack(op->g_out.ptr, op->g_sc.ptr, npix, xd, op->stream);

	if (!xdev)
		BAM_HIP(hipMemcpyAsync(x0, xd, sizeof(float2) * npix, hipMemcpyDeviceToHost, op->stream));

	BAM_HIP(hipStreamSynchronize(op->stream));
}

int bart_amd_warmnet_guess_dir(struct bart_amd_warmnet* _net, const struct linop_s* _op, float l2lambda, int N,
			       const long img_dims[], const float _Complex* b, const float _Complex* d,
			       float _Complex* x0)
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		WarmNet* net = (WarmNet*)_net;
		BAM_CHECK(nullptr != op && nullptr != net && nullptr != b && nullptr != x0,
			  "bart_amd_warmnet_guess: NULL argument");
		std::lock_guard<std::mutex> lock(op->mtx());
		if (0 != op->user_stream)
			BAM_HIP(hipStreamSynchronize(op->user_stream));
		warm_guess(net, op, l2lambda, N, img_dims, b, d, x0);
		return 0;
	}, -1);
}

// out[c M + m] = y[c M + m] w[m] (out may be y)
__global__ void k_kspace_weight(const float2* y, const float* __restrict__ w, long M, long n, float2* out)
{
	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
		out[e] = cscale(y[e], w[e % M]);
}

int bart_amd_warmnet_guess_dcf(struct bart_amd_warmnet* _net, const struct linop_s* _op, float l2lambda, int N,
			       const long img_dims[], const float _Complex* b, const float _Complex* ksp,
			       const float* wgt, long nwgt, float _Complex* x0)
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		WarmNet* net = (WarmNet*)_net;
		BAM_CHECK(nullptr != op && nullptr != net && nullptr != b && nullptr != ksp && nullptr != wgt &&
				  nullptr != x0,
			  "bart_amd_warmnet_guess_dcf: NULL argument");
		BAM_CHECK(nwgt > 0 && 0 == op->osize % nwgt, "bart_amd_warmnet_guess_dcf: %ld weights for %ld samples",
			  nwgt, op->osize);
		std::lock_guard<std::mutex> lock(op->mtx());
		if (0 != op->user_stream)
			BAM_HIP(hipStreamSynchronize(op->user_stream));
		if ((long)op->g_kw.n < op->osize)
			op->g_kw.alloc(op->osize);
		if ((long)op->g_w.n < nwgt)
			op->g_w.alloc(nwgt);
		if ((long)op->g_dcf.n < op->isize)
			op->g_dcf.alloc(op->isize);
		const float2* yd = (const float2*)ksp;
		if (!on_device(ksp)) {
			BAM_HIP(hipMemcpyAsync(op->g_kw.ptr, ksp, sizeof(float2) * op->osize, hipMemcpyHostToDevice,
					       op->stream));
			yd = op->g_kw.ptr;
		}
		BAM_HIP(hipMemcpyAsync(op->g_w.ptr, wgt, sizeof(float) * nwgt,
				       on_device(wgt) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, op->stream));
		const unsigned blocks = (unsigned)std::min<long>((op->osize + 255) / 256, 8192);
		hipLaunchKernelGGL(k_kspace_weight, dim3(blocks), dim3(256), 0, op->stream, yd, op->g_w.ptr, nwgt,
				   op->osize, op->g_kw.ptr);
		BAM_HIP(hipGetLastError());
		op->adj(op->g_dcf.ptr, op->g_kw.ptr);
		warm_guess(net, op, l2lambda, N, img_dims, b, (const float _Complex*)op->g_dcf.ptr, x0);
		return 0;
	}, -1);
}

int bart_amd_warmnet_guess(struct bart_amd_warmnet* net, const struct linop_s* op, float l2lambda, int N,
			   const long img_dims[], const float _Complex* b, float _Complex* x0)
{
	return bart_amd_warmnet_guess_dir(net, op, l2lambda, N, img_dims, b, nullptr, x0);
}

int bart_amd_tuning_report(char* buf, int n)
{
	return guarded([&]() -> int {
		const std::string r = tuning_report();
		if (nullptr != buf && n > 0) {
			strncpy(buf, r.c_str(), (size_t)n - 1);
			buf[n - 1] = 0;
		}
		return (int)r.size();
	}, -1);
}

// Strang's circulant preconditioner of a Toeplitz operator (opaque to C)
struct Precond {
	Op* op;
	DevBuf<float2> pre;
	int ncoil;
};

const struct operator_s* nufft_precond_create(const struct linop_s* _op)
{
	return guarded([&]() -> const struct operator_s* {
		Op* op = (Op*)_op;
		BAM_CHECK(nullptr != op && OP_NUDFT != op->kind, "nufft_precond_create: not a NUFFT operator");
		BAM_CHECK(op->toeplitz(), "nufft_precond_create: the operator must be in Toeplitz mode");
		BAM_CHECK(1 == op->nchunks(), "nufft_precond_create: not with the low-memory modes");
		BAM_CHECK(0 == op->K, "nufft_precond_create: not for an operator with a subspace basis");
		BAM_CHECK(op->frames.empty(), "nufft_precond_create: not for per-frame trajectories");
		std::lock_guard<std::mutex> lock(op->mtx());
		auto p = new Precond{ op, {}, (int)(op->isize / op->nufft->N) };
		try {
			op->nufft->compute_precond(p->pre);
		} catch (...) {
			delete p;
			throw;
		}
		return (const struct operator_s*)p;
	}, nullptr);
}

void operator_free(const struct operator_s* _p)
{
	guarded_void([&]() {
		const Precond* p = (const Precond*)_p;
		if (nullptr != p) {
			BAM_HIP(hipStreamSynchronize(p->op->stream));
			delete p;
		}
	});
}

int lsqr_conjgrad_precond(const struct linop_s* _op, const struct operator_s* _pre, int maxiter, float l2lambda,
			  float tol, int N, const long img_dims[], float _Complex* x, const float _Complex* b,
			  bool warmstart, float* resid_trace, struct cg_stats_s* stats)
{
	return guarded([&]() -> int {
		Op* op = (Op*)_op;
		const Precond* pc = (const Precond*)_pre;
		BAM_CHECK(nullptr != op && nullptr != pc && pc->op == op,
			  "lsqr_conjgrad_precond: the preconditioner belongs to another operator");
		check_dims("lsqr_conjgrad_precond", N, img_dims, op->N, op->idims);
		std::lock_guard<std::mutex> lock(op->mtx());
		if (0 != op->user_stream)
			BAM_HIP(hipStreamSynchronize(op->user_stream));
		op->set_plan_stream(op->stream);

		const long ne = op->isize;
		// its own solver: the preconditioned iteration is another graph
		CgSolver cg(2 * ne, op->stream);

		DevBuf<float2> xs(ne), bs(ne), bp(ne);
		BAM_HIP(hipMemcpyAsync(bs.ptr, b, sizeof(float2) * ne, hipMemcpyDefault, op->stream));
		if (warmstart)
			BAM_HIP(hipMemcpyAsync(xs.ptr, x, sizeof(float2) * ne, hipMemcpyDefault, op->stream));

		// CG on P (A^H A + lambda I) x = P b (iter/lsqr.c:126-142: the
		// preconditioner chained after the normal equations and the adjoint;
		// CG's own lambda 0)
		const float lam = l2lambda;
		op->nufft->apply_precond(pc->pre.ptr, bp.ptr, bs.ptr, pc->ncoil);
		NormalFn normal = [op, pc, lam, ne](float2* d, const float2* s, const int* skip, const CgApFuse*) {
			op->nrm(d, s, skip);
			axpy(d, s, lam, ne, op->nufft->stream, skip);
			op->nufft->apply_precond(pc->pre.ptr, d, d, pc->ncoil);
			return false;
		};

		CgResult r = cg.solve(normal, 0.f, maxiter, tol, xs.ptr, bp.ptr, warmstart, false, true);
		BAM_HIP(hipMemcpy(x, xs.ptr, sizeof(float2) * ne, hipMemcpyDefault));

		if (nullptr != resid_trace && r.iterations > 0)
			memcpy(resid_trace, r.trace.data(), sizeof(float) * r.iterations);

		if (nullptr != stats) {
			stats->iterations = r.iterations;
			stats->status = r.status;
			stats->resid = r.resid;
			stats->bnorm = r.bnorm;
			stats->time_ms = r.ms;
			stats->fell_back = false;
		}

		return r.iterations;
	}, -1);
}

int bart_amd_precond_apply(const struct operator_s* _pre, float _Complex* dst, const float _Complex* src)
{
	return guarded([&]() -> int {
		const Precond* pc = (const Precond*)_pre;
		BAM_CHECK(nullptr != pc, "bart_amd_precond_apply: NULL operator");
		Op* op = pc->op;
		std::lock_guard<std::mutex> lock(op->mtx());
		const long ne = op->isize;
		DevBuf<float2> d(ne);
		BAM_HIP(hipMemcpyAsync(d.ptr, src, sizeof(float2) * ne, hipMemcpyDefault, op->stream));
		op->nufft->apply_precond(pc->pre.ptr, d.ptr, d.ptr, pc->ncoil);
		BAM_HIP(hipMemcpyAsync(dst, d.ptr, sizeof(float2) * ne, hipMemcpyDefault, op->stream));
		BAM_HIP(hipStreamSynchronize(op->stream));
		return 0;
	}, -1);
}

// ---------------------------------------------------------------------------
// ESPIRiT (calib/calib.c:686-807, misc/mri2.c:171-224)
// ---------------------------------------------------------------------------

const struct ecalib_conf ecalib_defaults = {
	.kdims = { 6, 6, 6 },
	.threshold = 0.001f,
	.numsv = -1,
	.percentsv = -1.f,
	.weighting = false,
	.softcrop = false,
	.crop = 0.8f,
	.orthiter = true,
	.num_orthiter = 30,
	.usegpu = false,
	.perturb = -1.f,
	.intensity = false,
	.rotphase = true,
	.var = -1.f,
	.automate = false,
	.phase_normalize = false,
	.econdim = -2,
	.nystroem = false,
	.nystroem_os = 20,
	.nystroem_K = 40,
};

float _Complex* extract_calib(long caldims[16], const long calsize[3], const long in_dims[16],
			      const float _Complex* in_data, bool fixed)
{
	return guarded([&]() -> float _Complex* {
		BAM_CHECK(nullptr != in_data, "extract_calib: NULL data");
		BAM_CHECK(1 == in_dims[4], "MAPS dimension is not of size one.");
		for (int i = 5; i < 16; i++)
			BAM_CHECK(1 == in_dims[i], "extract_calib: dims >= 5 must be 1");
		BAM_CHECK(!fixed, "extract_calib: the fixed-size region is not supported");
		const int C = (int)in_dims[3];
		const long n = md_size(4, in_dims);
		std::vector<float2> h = to_host((const float2*)in_data, n);
		long cd[3];
		std::vector<float2> cal = extract_calib_host(calsize, in_dims, C, h.data(), cd);
		for (int i = 0; i < 16; i++)
			caldims[i] = (i < 3) ? cd[i] : ((3 == i) ? C : 1);
		float2* out = (float2*)malloc(sizeof(float2) * cal.size());
		BAM_CHECK(nullptr != out, "extract_calib: out of host memory");
		memcpy(out, cal.data(), sizeof(float2) * cal.size());
		return (float _Complex*)out;
	}, nullptr);
}

void calib(const struct ecalib_conf* conf, const long out_dims[16], float _Complex* out_data, float _Complex* eptr,
	   int SN, float svals[], const long calreg_dims[16], const float _Complex* data)
{
	guarded_void([&]() {
		BAM_CHECK(nullptr != conf && nullptr != out_data && nullptr != data, "calib: NULL argument");
		BAM_CHECK(!conf->weighting && !conf->automate && -1 == conf->numsv && -1.f == conf->percentsv,
			  "calib: soft weighting and automatic or count thresholds are not supported");
		BAM_CHECK(!conf->intensity && !conf->phase_normalize && !conf->nystroem && !conf->softcrop,
			  "calib: intensity and phase normalisation, Nystroem and soft crop are not supported");
		BAM_CHECK(conf->orthiter && conf->perturb <= 0.f, "calib: only the orthogonal iteration is supported");
		BAM_CHECK(conf->crop >= 0.f, "calib: the automatic crop threshold is not supported");
		for (int i = 5; i < 16; i++)
			BAM_CHECK(1 == out_dims[i] && 1 == calreg_dims[i], "calib: dims >= 5 must be 1");
		BAM_CHECK(out_dims[3] == calreg_dims[3], "calib: coil dims differ");
		const int C = (int)out_dims[3];
		const int M = (int)out_dims[4];
		const long N = conf->kdims[0] * conf->kdims[1] * conf->kdims[2] * C;
		BAM_CHECK(nullptr == svals || SN >= (int)std::min<long>(N, 1 << 30), "calib: svals too short");

		EspiritConf ec;
		for (int d = 0; d < 3; d++)
			ec.kdims[d] = (1 == out_dims[d]) ? 1 : conf->kdims[d];
		ec.threshold = conf->threshold;
		ec.crop = conf->crop;
		ec.softcrop = conf->softcrop;
		ec.num_orthiter = conf->num_orthiter;
		ec.rotphase = conf->rotphase;

		const long img[3] = { out_dims[0], out_dims[1], out_dims[2] };
		const long cd[3] = { calreg_dims[0], calreg_dims[1], calreg_dims[2] };
		std::vector<float2> cal = to_host((const float2*)data, md_size(4, calreg_dims));

		const long np = img[0] * img[1] * img[2];
		hipStream_t st = default_stream();
		DevBuf<float2> dm((size_t)M * C * np), de((size_t)M * np);
		std::vector<float> sv(N);
		espirit_calib(ec, img, C, M, cd, cal.data(), dm.ptr, de.ptr, sv.data(), st);

		BAM_HIP(hipMemcpy(out_data, dm.ptr, sizeof(float2) * M * C * np, hipMemcpyDefault));
		if (nullptr != eptr)
			BAM_HIP(hipMemcpy(eptr, de.ptr, sizeof(float2) * M * np, hipMemcpyDefault));
		if (nullptr != svals)
			for (long i = 0; i < N && i < SN; i++)
				svals[i] = sv[i];
	});
}

}  // extern "C"
