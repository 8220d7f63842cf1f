/*
 * bart_amd.h -- C ABI of the gfx950 NUFFT / CG-SENSE library.
 *
 * The entry points keep the names, argument order and meaning of BART's
 * interfaces for this path so that a BART build (or any C/FFI caller) can
 * link them in place of the CPU/CUDA implementation.  Array arguments use
 * BART's layout: N dims (normally 16, misc/mri.h:9-27), column-major,
 * complex float (re, im interleaved).  Pointers may be host or device
 * (hipMalloc) memory; the library checks where each pointer lives, as BART
 * selects its GPU path by cuda_ondevice() (num/gpuops.c:445).  Host pointers
 * are staged through device buffers owned by the operator.
 *
 * Errors follow BART's error() (misc/misc.c:95-128): by default a message is
 * printed and the process exits with status 1.  After
 * bart_amd_set_error_mode(BART_AMD_ERRORS_RETURN) the message is kept
 * (bart_amd_last_error()) and the call returns -- constructors return NULL.
 */
#ifndef BART_AMD_H
#define BART_AMD_H

#include <complex.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#define BAM_COMPLEX float _Complex
#else
#define BAM_COMPLEX _Complex float
#endif

struct linop_s;

/* noncart/nufft.h:7-29 (field order kept: the struct is passed by value) */
struct nufft_conf_s {

	bool toeplitz;
	bool pcycle;
	bool periodic;
	bool lowmem;
	unsigned long flags;
	unsigned long cfft;
	bool decomp;
	bool nopsf;
	bool upper_triag;
	bool real;
	bool compress_psf;
	bool decomposed_psf;

	bool precomp;
	bool precomp_linphase;
	bool precomp_fftmod;
	bool precomp_roll;
	bool zero_overhead;

	float width;
	float os;
};

/* noncart/nufft.h:32 (values of noncart/nufft.c:53-75) */
extern struct nufft_conf_s nufft_conf_defaults;

/* noncart/nufft.h:48-57 -- replaces nufft_create2 (noncart/nufft.c:1174-1276).
 * Supported: decomp = true, kernel width <= 14 (<= 15 with os != 2), 1
 * to 3 transformed dims (flags = 7), samples in dims 1-2 and coils in dim
 * 3.  os != 2 gives the chained NUFFT of noncart/nufft_chain.c:192-244
 * (round(os n) grid, no Toeplitz operator, no basis).
 * A subspace basis (bas_dims [1 1 1 1 1 T K], K <= 16,
 * noncart/nufft.c:1018-1036) takes coil images [X Y Z C 1 1 K], k-space
 * [1 R S C 1 T] (dim 6 1 or K; the operator's output has 1) and a
 * trajectory and weights with dim 5 of 1 or T; its Toeplitz normal operator
 * uses the K x K PSF.  Modes that only change memory use (lowmem,
 * zero_overhead, precomp*, upper_triag, pcycle = false) give the same
 * results and are accepted; compress_psf is rejected. */
struct linop_s* nufft_create2(int N, const long ksp_dims[], const long cim_dims[], const long traj_dims[],
			      const BAM_COMPLEX* traj, const long wgh_dims[], const BAM_COMPLEX* weights,
			      const long bas_dims[], const BAM_COMPLEX* basis, struct nufft_conf_s conf);

/* noncart/nufft.h:40-46 -- replaces nufft_create (noncart/nufft.c:1315-1327) */
struct linop_s* nufft_create(int N, const long ksp_dims[], const long cim_dims[], const long traj_dims[],
			     const BAM_COMPLEX* traj, const BAM_COMPLEX* weights, struct nufft_conf_s conf);

/* sense/modelnc.h -- replaces sense_nc_init (sense/modelnc.c:28-67): the
 * linear operator image -> k-space  A = NUFFT S  with normal operator
 * S^H T S (Toeplitz) when conf->toeplitz.  One set of maps (MAPS dim 1).
 * With a basis, max_dims holds the K coefficients in dim 6 (and 1 in dim
 * 5), as pics sets it (src/pics.c:326-333).
 * When fft_opp is not NULL it receives the NUFFT operator (coil images <->
 * k-space) sharing the same plan and PSF; free it with linop_free. */
const struct linop_s* sense_nc_init(const long max_dims[16], const long map_dims[16], const BAM_COMPLEX* maps,
				    const long ksp_dims[16], const long traj_dims[16], const BAM_COMPLEX* traj,
				    const struct nufft_conf_s* conf, const long wgs_dims[16], const BAM_COMPLEX* weights,
				    const long basis_dims[16], const BAM_COMPLEX* basis,
				    const struct linop_s** fft_opp, unsigned long shared_img_dims);

/* noncart/nudft.h -- replaces nudft_create (noncart/nudft.c:283-348): the
 * direct non-uniform DFT (nufft -s), ksp[c][m] = s sum_x img[c][x]
 * exp(-2 pi i k_m (x - n/2) / n), s = 1 / sqrt(size of the coil-image
 * array) as BART.  flags = 7; no B0 term (fieldmap and timemap NULL).
 * Its normal operator is adjoint(forward(x)). */
const struct linop_s* nudft_create(int N, unsigned long flags, const long odims[], const long idims[],
				   const long tdims[], const BAM_COMPLEX* traj, const long fmdims[],
				   const BAM_COMPLEX* fieldmap, const long tmdims[], const BAM_COMPLEX* timemap);

/* linops/linop.h:66-72 -- replace linop_forward/adjoint/normal
 * (linops/linop.c:406-525).  dst is overwritten. */
void linop_forward(const struct linop_s* op, int DN, const long ddims[], BAM_COMPLEX* dst, int SN,
		   const long sdims[], const BAM_COMPLEX* src);
void linop_adjoint(const struct linop_s* op, int DN, const long ddims[], BAM_COMPLEX* dst, int SN,
		   const long sdims[], const BAM_COMPLEX* src);
void linop_normal(const struct linop_s* op, int N, const long dims[], BAM_COMPLEX* dst, const BAM_COMPLEX* src);

/* linops/linop.h:55 -- replaces linop_free (linops/linop.c:998) */
void linop_free(const struct linop_s* op);

/* noncart/nufft.h:92-99 -- replace nufft_get_psf_dims / nufft_get_psf2 /
 * nufft_get_psf / nufft_update_psf2 / nufft_update_psf
 * (noncart/nufft.c:2008-2073); PSF dims are N + 1 with the decomposition
 * index last; psf_strs are byte strides (BART's md strides), and psf may be
 * host or device memory. */
int nufft_get_psf_dims(const struct linop_s* nufft, int N, long psf_dims[]);
void nufft_get_psf2(const struct linop_s* nufft, int N, const long psf_dims[], const long psf_strs[],
		    BAM_COMPLEX* psf);
void nufft_get_psf(const struct linop_s* nufft, int N, const long psf_dims[], BAM_COMPLEX* psf);
void nufft_update_psf2(const struct linop_s* nufft, int ND, const long psf_dims[], const long psf_strs[],
		       const BAM_COMPLEX* psf);
void nufft_update_psf(const struct linop_s* nufft, int ND, const long psf_dims[], const BAM_COMPLEX* psf);

/* noncart/nufft.h:84-86 -- replaces nufft_create_normal
 * (noncart/nufft.c:1285-1312): the Toeplitz normal operator of a PSF alone,
 * on coil images cim_dims; psf (dims N + 1, may be NULL and set later with
 * nufft_update_psf) is copied.  Its forward and adjoint are errors ("NuFFT
 * with normal operator only!"); linop_normal applies it.  basis must be
 * false. */
struct linop_s* nufft_create_normal(int N, const long cim_dims[], int ND, const long psf_dims[],
				    const BAM_COMPLEX* psf, bool basis, struct nufft_conf_s conf);

/* noncart/nufft.h:88-91 -- replaces nufft_update_traj
 * (noncart/nufft.c:2041-2054, nufft_set_traj :1002-1159): a new trajectory
 * (and weights) of the same sample count for an existing NUFFT or SENSE
 * operator; the gridding plan and, in Toeplitz mode, the PSF are rebuilt.
 * basis must be NULL. */
void nufft_update_traj(const struct linop_s* nufft, int N, const long trj_dims[], const BAM_COMPLEX* traj,
		       const long wgh_dims[], const BAM_COMPLEX* weights, const long bas_dims[],
		       const BAM_COMPLEX* basis);

/* noncart/nufft.h:70-73 -- replaces compute_psf2 (noncart/nufft.c:727-776):
 * the decomposed PSF of a trajectory, psf_dims = image dims, 1..., 2^d in
 * dim N.  flags = 7, basis NULL, upper_triag false; lowmem changes nothing.
 * The result is allocated where the trajectory lives, as BART's
 * md_alloc_sameplace: malloc() host memory for a host trajectory (release
 * with free()), hipMalloc() device memory for a device trajectory (release
 * with hipFree()).  NULL on error. */
BAM_COMPLEX* compute_psf2(int N, const long psf_dims[], unsigned long flags, const long trj_dims[],
			  const BAM_COMPLEX* traj, const long bas_dims[], const BAM_COMPLEX* basis,
			  const long wgh_dims[], const BAM_COMPLEX* weights, bool periodic, bool lowmem,
			  bool upper_triag);

/* The stream an operator's applies launch on (BART runs each OpenMP
 * thread's GPU work on a stream of its own, num/gpuops.c:214-278): a
 * hipStream_t, or NULL for the handle's own stream.  With device pointers
 * and a stream set, linop_forward / adjoint / normal enqueue on it and
 * return without waiting; lsqr_conjgrad waits for the stream before it
 * reads x and b.  Handles that share a plan (sense_nc_init's fft_opp)
 * share the lock, not the stream setting.  Returns 0, or -1 on error. */
int bart_amd_linop_set_stream(const struct linop_s* op, void* stream);

/* num/fft.h -- replace fft/ifft (num/fft.c:387-417): unnormalised FFT,
 * sign -1 (fft) / +1 (ifft) over the dims in flags (subset of dims 0-2);
 * dims above 2 are batch.  dst may equal src. */
void fft(int D, const long dimensions[], unsigned long flags, BAM_COMPLEX* dst, const BAM_COMPLEX* src);
void ifft(int D, const long dimensions[], unsigned long flags, BAM_COMPLEX* dst, const BAM_COMPLEX* src);

/* Conjugate gradients on (A^H A + l2lambda I) x = b with A^H A the normal
 * operator of `op` -- Toeplitz when the operator was created with
 * conf.toeplitz, else adjoint(forward(x)) (BART's lsqr2 -> iter2_conjgrad -> conjgrad chain:
 * iter/lsqr.c:52-170, iter/iter2.c:227-270, iter/italgos.c:497-580).
 * Stops after maxiter iterations or when ||r|| <= tol ||b|| (tol = 0 as in
 * BART: fixed iteration count).  x holds the initial guess when warmstart
 * (pics -W, iter/itop.c:54-66), else zero is used.  With warm_fallback the
 * zero start is used whenever the warm start's residual exceeds ||b||.
 * Returns the number of iterations, or -1 on error. */
struct cg_stats_s {

	int iterations;
	int status;        /* 1 maxiter, 2 converged, 3 pAp = 0, 4 nothing to do */
	float resid;       /* ||r|| at exit */
	float bnorm;       /* ||b|| */
	float time_ms;     /* device time of the solve */
	bool fell_back;
};

int lsqr_conjgrad(const struct linop_s* op, int maxiter, float l2lambda, float tol, int N, const long img_dims[],
		  BAM_COMPLEX* x, const BAM_COMPLEX* b, bool warmstart, bool warm_fallback, float* resid_trace,
		  struct cg_stats_s* stats);

/* The autotuner's choices (no counterpart in BART): the adjoint gridding
 * launch of every NUFFT plan is tuned when the plan is built -- coils per
 * work-group and persistent work-groups per compute unit, timed on the plan
 * (results do not depend on them) -- and kept per shape class in a cache file
 * (BART_AMD_TUNE_FILE; BART_AMD_AUTOTUNE=off keeps the defaults).  Writes a
 * JSON list of this process's choices into buf (n bytes, truncated) and
 * returns its full length, -1 on error. */
int bart_amd_tuning_report(char* buf, int n);

/* iter/monitor.c (iter_monitor_create) as pics -T uses it (src/pics.c:548-575):
 * a truth image for the CG solves (lsqr_conjgrad) of op.  Each later solve
 * writes err[k] = ||truth - x_k|| / ||truth|| for the iterate x_k that
 * iteration k starts from, as conjgrad calls its monitor
 * (iter/italgos.c:534); err is a host array of at least maxiter floats that
 * stays valid while the monitor is set; truth NULL ends the monitoring.
 * truth: host or device image of op's image dims.  Returns 0, -1 on error. */
int bart_amd_cg_monitor(const struct linop_s* op, const BAM_COMPLEX* truth, float* err);

/* lsqr_conjgrad without any device allocation: no counterpart in BART.  For
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same solve, with the same
 * results bit for bit, but r, p, A p and the solver state live in a static
 * working array of the library (one per GPU, 48 MiB + 64 KiB, held until the
 * process ends), the iterations are launched one by one (no graph) and the
 * stop flag is read after every 8.  Limits: at most 2^21 complex image
 * elements (e.g. 128^3 or 1448^2) and 4096 iterations -- larger requests
 * fail with a message; x and b must be device pointers; an operator without
 * conf.toeplitz needs its k-space buffer from an earlier lsqr_conjgrad call.
 * Every call on a GPU shares that GPU's working array: calls that overlap on
 * one GPU race. */
int lsqr_conjgrad_noalloc(const struct linop_s* op, int maxiter, float l2lambda, float tol, int N,
			  const long img_dims[], BAM_COMPLEX* x, const BAM_COMPLEX* b, bool warmstart,
			  bool warm_fallback, float* resid_trace, struct cg_stats_s* stats);

/* noncart/precond.h -- replaces nufft_precond_create (noncart/precond.c:127-164):
 * Strang's circulant preconditioner of a Toeplitz NUFFT (or SENSE) operator,
 * pre = |fftuc(cropped PSF)| + 1e-3 (precond.c:46-71), applied as
 * dst = IFFT(FFT(src) / pre) with unnormalised FFTs over the image dims
 * (precond.c:92-106).  Released with operator_free. */
struct operator_s;
const struct operator_s* nufft_precond_create(const struct linop_s* nufft_op);
void operator_free(const struct operator_s* op);
/* the preconditioner on one vector of the operator's image dims (host or
 * device); returns 0, -1 on error */
int bart_amd_precond_apply(const struct operator_s* precond, BAM_COMPLEX* dst, const BAM_COMPLEX* src);
/* lsqr2 with a preconditioner (iter/lsqr.c:126-142, as nufft -i -c uses it):
 * CG on P (A^H A + l2lambda I) x = P b, P = precond, CG's own lambda 0; the
 * stop rule of lsqr_conjgrad on the preconditioned residual.  x, b host or
 * device vectors.  Returns the number of iterations, -1 on error. */
int lsqr_conjgrad_precond(const struct linop_s* op, const struct operator_s* precond, int maxiter, float l2lambda,
			  float tol, int N, const long img_dims[], BAM_COMPLEX* x, const BAM_COMPLEX* b, bool warmstart,
			  float* resid_trace, struct cg_stats_s* stats);

/* Batched CG: nprob independent problems (A_i^H A_i + l2lambda I) x_i = b_i,
 * each with its own operator ops[i] (image size, coils and trajectory may
 * differ), solved together on the device.  Replaces BART's tool-level loop
 * over slices (`bart -l/-p ... pics`, src/bart.c:200-330) around lsqr2; the
 * vector updates of all problems are one launch each with per-problem
 * scalars, as conjgrad_batch (iter/italgos.c:589-680) does, and every
 * problem stops by the rule of lsqr_conjgrad, so x_i, the iteration count
 * and the residual trace of problem i equal those of lsqr_conjgrad on it.
 * The problems are spread over several streams; one iteration is captured
 * as a hipGraph per stream.  No two problems may share a NUFFT plan.  x[i],
 * b[i]: host or device vectors of ops[i]'s image size (x[i] holds the warm
 * start when warmstart).  resid_trace: NULL or nprob rows of max(1, maxiter)
 * floats.  stats: NULL or nprob entries; time_ms is the device time of the
 * stream that solved the problem.  Returns the largest iteration count, or
 * -1 on error. */
int lsqr_conjgrad_batch(int nprob, const struct linop_s* const ops[], int maxiter, float l2lambda, float tol, int N,
			BAM_COMPLEX* const x[], const BAM_COMPLEX* const b[], bool warmstart, bool warm_fallback,
			float* resid_trace, struct cg_stats_s stats[]);

/* Layers of the CG warm-start network: no counterpart in BART (the network
 * that proposes x0 for pics -W, iter/itop.c:54-66, is new).  3x3
 * convolution, zero padding, cross-correlation, on the matrix cores with
 * bfloat16 operands (rounded to nearest even) and FP32 accumulation:
 *   y[b][co][i][j] = bias[co] + sum_{ci,u,v} w[co][ci][u][v] x[b][ci][i+u-1][j+v-1]
 * Device pointers only; activations [B][C][H][W], weights [Cout][Cin][3][3],
 * float; 1 to 64 channels.  Launched on the null stream, not synchronised.
 * Return 0, or -1 on error. */
/* y = [relu](conv(x, w) + bias) [+ res]; bias and res may be NULL */
int bart_amd_conv3x3(const float* x, int B, int Cin, int H, int W, const float* w, const float* bias,
		     const float* res, int Cout, int relu, float* y);
/* dx = gradient with respect to x for the output gradient dy */
int bart_amd_conv3x3_bwd_data(const float* dy, int B, int Cout, int H, int W, const float* w, int Cin, float* dx);
/* dw, db = gradients with respect to w and bias (db may be NULL); W a
 * multiple of 8, at most 256 */
int bart_amd_conv3x3_bwd_weights(const float* x, const float* dy, int B, int Cin, int Cout, int H, int W, float* dw,
				 float* db);

/* The warm-start network (no counterpart in BART): six of the layers above,
 * 2 -> channels (x5) -> 2, ReLU after the first five, the input added to the
 * output; trained on the mean squared error with Adam (lr given, betas 0.9
 * and 0.999, eps 1e-8).  x, y, target: device [B][2][H][W] (real and
 * imaginary parts of a complex image).  Parameters are packed
 * [W0 b0 W1 b1 ... W5 b5]; creation draws He-normal weights from the seed. */
struct bart_amd_warmnet;
struct bart_amd_warmnet* bart_amd_warmnet_create(int channels, unsigned long seed);
void bart_amd_warmnet_free(struct bart_amd_warmnet* net);
long bart_amd_warmnet_num_params(const struct bart_amd_warmnet* net);
/* host buffer of num_params floats: set = 0 reads the parameters, 1 writes
 * them, 2 reads the gradients of the last training step */
int bart_amd_warmnet_params(struct bart_amd_warmnet* net, float* host, int set);
int bart_amd_warmnet_forward(struct bart_amd_warmnet* net, const float* x, int B, int H, int W, float* y);
/* one Adam step; returns the loss of x before the step, or -1 on error */
double bart_amd_warmnet_train_step(struct bart_amd_warmnet* net, const float* x, const float* target, int B, int H,
				   int W, float lr);

/* one Adam step for a loss whose gradient with respect to the output y of
 * the last bart_amd_warmnet_forward(net, x, ...) is dout ([B][2][H][W]),
 * e.g. the gradient of the normal-equation residual of the guess; x must be
 * that forward's input */
int bart_amd_warmnet_train_step_grad(struct bart_amd_warmnet* net, const float* x, const float* dout, int B, int H,
				     int W, float lr);

/* The network's input and the learned initial guess, computed on the
 * device (no counterpart in BART; the x0 of pics -W, src/pics.c:580-594,
 * iter/itop.c:54-66).  x1 = alpha b is the first CG iterate from zero on
 * (A^H A + l2lambda I) x = b (alpha = ||b||^2 / b.(A^H A b + l2lambda b),
 * as CG computes it; A^H A the operator's normal operator), s = max |x1|
 * its scale, and the input channels [2][Y][X] (real, imaginary) of x1 / s.
 * 2-D images of op's image dims; b, x0 host or device vectors; inp, out
 * device arrays.  Each returns 0, -1 on error. */
/* inp = channels of x1 / s, *scale = s */
int bart_amd_warmnet_input(const struct linop_s* op, float l2lambda, int N, const long img_dims[],
			   const BAM_COMPLEX* b, float* inp, float* scale);
/* out = channels of img / scale (a training target; img and out device) */
int bart_amd_warmnet_pack(int N, const long img_dims[], const BAM_COMPLEX* img, float scale, float* out);
/* x0 = s (x1 / s + net(x1 / s)) */
int bart_amd_warmnet_guess(struct bart_amd_warmnet* net, const struct linop_s* op, float l2lambda, int N,
			   const long img_dims[], const BAM_COMPLEX* b, BAM_COMPLEX* x0);
/* The same two along a direction d instead of b (no counterpart in BART):
 * x1 = alpha d with alpha = d^H b / d^H (A^H A d + l2lambda d), the best
 * multiple of d for the residual of (A^H A + l2lambda I) x = b; for d the
 * density-compensated adjoint A^H (w y) this is the network input of the
 * learned warm start.  d = NULL means d = b. */
int bart_amd_warmnet_input_dir(const struct linop_s* op, float l2lambda, int N, const long img_dims[],
			       const BAM_COMPLEX* b, const BAM_COMPLEX* d, float* inp, float* scale);
int bart_amd_warmnet_guess_dir(struct bart_amd_warmnet* net, const struct linop_s* op, float l2lambda, int N,
			       const long img_dims[], const BAM_COMPLEX* b, const BAM_COMPLEX* d, BAM_COMPLEX* x0);
/* The guess along the density-compensated adjoint d = A^H (w y), computed
 * in the library: ksp = y (op's k-space, host or device), wgt = w, nwgt
 * weights repeated over the k-space (one per sample of a coil; nwgt must
 * divide the k-space size).  No counterpart in BART. */
int bart_amd_warmnet_guess_dcf(struct bart_amd_warmnet* net, const struct linop_s* op, float l2lambda, int N,
			       const long img_dims[], const BAM_COMPLEX* b, const BAM_COMPLEX* ksp, const float* wgt,
			       long nwgt, BAM_COMPLEX* x0);

/* Error handling (BART's error()/catcher semantics, misc/misc.c:95-128). */
enum { BART_AMD_ERRORS_EXIT = 0, BART_AMD_ERRORS_RETURN = 1 };
void bart_amd_set_error_mode(int mode);
const char* bart_amd_last_error(void);
void bart_amd_clear_error(void);

/* Per-kernel timing with HIP events on the launching stream (no counterpart
 * in BART): bart_amd_profile(1) starts recording, the report returns
 * {"kernel": [count, total_ms, min_ms, max_ms], ...} as JSON and returns the
 * length needed; the records are cleared only when the report fits into buf.
 * Launches inside stream capture are not recorded. */
void bart_amd_profile(int enable);
int bart_amd_profile_report(char* buf, int len);

/* Library version string. */
const char* bart_amd_version(void);

/* calib/calib.h:12-34 -- struct ecalib_conf, field for field */
struct ecalib_conf {
	long kdims[3];
	float threshold;
	int numsv;
	float percentsv;
	bool weighting;
	bool softcrop;
	float crop;
	bool orthiter;
	int num_orthiter;
	bool usegpu;
	float perturb;
	bool intensity;
	bool rotphase;
	float var;
	bool automate;
	bool phase_normalize;
	int econdim;
	bool nystroem;
	long nystroem_os;
	long nystroem_K;
};

/* calib/calib.h:36 (values of calib/calib.c:686-708) */
extern const struct ecalib_conf ecalib_defaults;

/* misc/mri2.h -- replaces extract_calib (misc/mri2.c:171-224): the
 * calibration region of Cartesian k-space [X Y Z C] (host or device): the
 * centred box that grows while fully sampled up to calsize, then moved along
 * the readout to maximal energy.  Returns a malloc'ed host array of
 * caldims; fixed = true is not supported. */
BAM_COMPLEX* extract_calib(long caldims[16], const long calsize[3], const long in_dims[16],
			   const BAM_COMPLEX* in_data, bool fixed);

/* calib/calib.h:38-39 -- replaces calib (calib/calib.c:801-807, calib2
 * :715-798): ESPIRiT maps out [X Y Z C M] and eigenvalue maps eptr
 * [X Y Z 1 M] (may be NULL) from the calibration region [cx cy cz C]; SN
 * singular values into svals (may be NULL).  Pointers host or device.  The
 * Gram matrix, its eigenvectors (rocSOLVER), the image-space covariance, its
 * sinc interpolation and the point-wise orthogonal iteration run on the GPU.
 * Supported: threshold and crop (hard), orthiter, rotphase; weighting,
 * automatic thresholds, intensity / phase normalisation, Nystroem and soft
 * crop are rejected.  The phase reference (the channel principal component)
 * has its first entry made real, so maps agree with BART's up to one global
 * phase per map. */
void calib(const struct ecalib_conf* conf, const long out_dims[16], BAM_COMPLEX* out_data, BAM_COMPLEX* eptr,
	   int SN, float svals[], const long calreg_dims[16], const BAM_COMPLEX* data);

#ifdef __cplusplus
}
#endif

#undef BAM_COMPLEX
#endif /* BART_AMD_H */
