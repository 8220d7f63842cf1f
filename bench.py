"""Benchmark of the NUFFT + CG-SENSE hot path on one MI355X.

Default workload (one GPU): 3-D radial kooshball, 128^3 image, 16 coils,
8192 phyllotaxis spokes x 256 samples (2,097,152 samples per coil), 2x
oversampled grid (decomposed into 8 sub-grids of 128^3), Kaiser-Bessel width 6.
One step = adjoint NUFFT (k-space -> coil images), forward NUFFT (coil images ->
k-space) and one CG-SENSE solve (A^H y, then 20 Toeplitz iterations,
lambda = 0.01).  Measured beside the steps in the same run: the kooshball
CG-SENSE solve to relative residual 1e-4, and the 2-D radial workload's
(256^2, 8 coils, 201 x 512) CG-SENSE solve to relative residual 1e-4 from
zero and from the learned warm start.

Prints one JSON line.  `value` is the NUFFT adjoint+forward throughput in
million coil-samples per second (2 C M per adj+fwd pair / time of the pair).
Inputs are resident in HBM before the timed region.  `roofline` is for the
kernel with the largest share of the timed region, timed live with HIP
events on the stream it runs on, charged with the compulsory bytes of the
unit of work it belongs to (one Toeplitz CG iteration: 8 N (9 + C + 2^d);
one NUFFT: 8 C M + 4 d M + 8 C N); its own input/output bytes are
reported beside.  `cpu_baseline` is the CPU oracle (a C/OpenMP restatement
of BART's CPU algorithm) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on this workload (each pass in a run of its own), FETCH_SIZE
# scaled by 2 and WRITE_SIZE by 1: the factors at which the same counters
# read 1 GiB for a 1 GiB device copy.
TRAFFIC = {
    ("3d_radial_kooshball_128^3_16coils_8192x256", "fft_toeplitz"): 3_222_000_000,
    ("3d_radial_kooshball_128^3_16coils_8192x256", "grid_adjoint"): 4_277_000_000,
    ("3d_radial_kooshball_128^3_16coils_8192x256", "grid_forward"): 12_900_000_000,
}


def dims16(*d):
    return list(d) + [1] * (16 - len(d))


def workload(name):
    from bart_amd import inputs as I
    return {"kooshball": I.config_kooshball, "radial2d": I.config_radial2d, "spiral": I.config_spiral}[name]()


def survey_bytes(w):
    """Compulsory HBM bytes of the units of work of the metric (each input
    read once, each output written once; complex64 8 B, coordinates 4 B per
    dim): one NUFFT application 8 C M + 4 d M + 8 C N, one Toeplitz CG
    iteration 8 N (9 + C + 2^d) (image in and out, maps, PSF, and the CG
    update reading x, r, p, Ap and writing x, r, p), one SENSE A^H y
    8 C M + 4 d M + 8 N + 8 C N."""
    X, Y, Z = w["img"]
    N = X * Y * Z
    C = w["C"]
    M = w["R"] * w["S"]
    d = sum(1 for v in w["img"] if v > 1)
    return dict(nufft=8 * C * M + 4 * d * M + 8 * C * N, iteration=8 * N * (9 + C + 2 ** d),
                sense_adjoint=8 * C * M + 4 * d * M + 8 * N + 8 * C * N)


# the unit of work each kernel class belongs to: a kernel that runs once
# per Toeplitz CG iteration is charged with the whole iteration's
# compulsory bytes, a kernel of the NUFFT with the whole application's
UNIT_OF = {
    "fft_toeplitz": "iteration", "fft_toeplitz_slice_fwd": "iteration", "fft_toeplitz_slice_inv": "iteration",
    "sum_coil_groups": "iteration", "fft_toeplitz_fwd": "iteration", "fft_toeplitz_inv": "iteration",
    "grid_adjoint": "nufft", "grid_forward": "nufft", "permute_ksp": "nufft", "fft_fwd_axis": "nufft",
    "fft_adj_axis": "nufft", "fft_fwd_first": "nufft", "fft_adj_last": "nufft", "fft_fwd_slice": "nufft",
    "fft_adj_slice": "nufft",
}


def algorithmic_bytes(kernel, w):
    """Compulsory HBM bytes of one launch of a kernel class: each input read
    once, each output written once (complex64 = 8 B, coordinates 4 B/dim).
    The NUFFT operator of the step works on coil images (no maps); the
    Toeplitz classes belong to the SENSE normal operator (maps read, one
    image in and out, coil-group partial images).  Where the x-y slices fit
    in LDS and the z length has a pair kernel, the Toeplitz stack holds one
    x-y transform per pair of sub-grids that differ only in the z shift
    (S/2 volumes per coil), and the z-pass reads the PSF of all S."""
    X, Y, Z = w["img"]
    N = X * Y * Z
    C = w["C"]
    M = w["R"] * w["S"]
    d = sum(1 for v in w["img"] if v > 1)
    S = 2 ** d
    grid = 8 * S * C * N
    G = w.get("groups", 1)
    pair = d == 3 and X == Y and X in (32, 64, 96, 128) and Z in (64, 128, 256, 512)
    tgrid = grid // 2 if pair else grid
    table = {
        # k-space (x conj(weights)) -> sorted, coil-interleaved copy
        "permute_ksp": 8 * C * M + 8 * C * M,
        # sorted k-space + coordinates in, sub-grid stack out
        "grid_adjoint": 8 * C * M + 4 * d * M + grid,
        # sub-grid stack + coordinates in, k-space out
        "grid_forward": grid + 4 * d * M + 8 * C * M,
        # one axis of the sub-grid stack: read + write (z-pass with pairs:
        # the pair stack on one side)
        "fft_fwd_axis": tgrid + grid,
        "fft_adj_axis": grid + tgrid,
        # coil images in, sub-grid stack out / back
        "fft_fwd_first": 8 * C * N + grid,
        "fft_adj_last": grid + 8 * C * N,
        # (3-D slice path with pairs: the slice kernels store / read one
        # x-y transform per pair, the z-pass writes / reads all S sub-grids)
        "fft_fwd_slice": 8 * C * N + tgrid,
        "fft_adj_slice": tgrid + 8 * C * N,
        # SENSE normal operator: image + maps in, stack out
        "fft_toeplitz_slice_fwd": 8 * N + 8 * C * N + tgrid,
        # z axis forward, x PSF, inverse: stack read + write, PSF read
        "fft_toeplitz": 2 * tgrid + 8 * S * N,
        # stack + maps in, coil-group partial images out
        "fft_toeplitz_slice_inv": tgrid + 8 * C * N + 8 * G * N,
        "sum_coil_groups": 8 * G * N + 8 * N,
    }
    return table.get(kernel)


def fp64_nudft_check(w, cim, ksp, nsub):
    """Relative L2 error of the GPU NUFFT forward (coil 0, the step's last
    output) against the oracle's FP64 direct NUDFT on nsub samples spread
    evenly over the trajectory."""
    import oracle as O
    M = w["R"] * w["S"]
    X, Y, Z = w["img"]
    N = X * Y * Z
    idx = np.linspace(0, M - 1, nsub).astype(np.int64)
    img0 = cim[:N].cpu().numpy().astype(np.complex128).reshape(Z, Y, X)
    chunk = max(1, int(2e8 // (16 * N)))
    ref = O.nudft_forward(w["img"], w["traj"][idx], img0, chunk=chunk)[0]
    got = ksp[:M].cpu().numpy()[idx].astype(np.complex128)
    # samples whose kernel lies inside the grid (non-periodic gridding, as
    # BART's default, drops the taps beyond the grid edge; the NUDFT does not)
    k = np.abs(np.asarray(w["traj"])[idx].real)
    lim = np.array([X, Y, Z]) / 2.0 - 2.0
    inner = np.all((k <= lim) | (np.array([X, Y, Z]) == 1), axis=1)
    e_in = float(np.linalg.norm(got[inner] - ref[inner]) / np.linalg.norm(ref[inner])) if inner.any() else None
    return dict(forward_relerr=float(np.linalg.norm(got - ref) / np.linalg.norm(ref)), samples=int(nsub), coil=0,
                forward_relerr_interior=e_in, interior_samples=int(inner.sum()))


def explicit_vs_toeplitz(w, x, sense_t):
    """SENSE normal operator: explicit adjoint(forward(x)) (grid / gridH) vs
    the Toeplitz PSF operator -- times and relative difference."""
    import torch
    import bart_amd as ba
    X, Y, Z = w["img"]
    C, R, S = w["C"], w["R"], w["S"]
    conf = ba.NufftConf.defaults(toeplitz=False)
    sense_e = ba.sense_nc_init(dims16(X, Y, Z, C), dims16(X, Y, Z, C), w["maps"], dims16(1, R, S, C),
                               dims16(3, R, S), ba.traj_array(w["traj"]), conf=conf)
    out = {}
    res = {}
    for name, op in (("explicit", sense_e), ("toeplitz", sense_t)):
        o = op.normal(x)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            op.normal(x, out=o)
        e1.record()
        torch.cuda.synchronize()
        out[f"{name}_normal_ms"] = round(e0.elapsed_time(e1) / 10, 4)
        res[name] = o
    out["toeplitz_vs_explicit_relerr"] = float((res["toeplitz"] - res["explicit"]).norm() / res["explicit"].norm())
    sense_e.free()
    return out


def cpu_probe():
    """Whether BART itself or FFTW (single precision) is installed on this
    host: the CPU baseline would then be BART's own path.  Neither can ship
    with this repository."""
    import ctypes.util
    import shutil
    return dict(bart=shutil.which("bart"), fftw3f=ctypes.util.find_library("fftw3f"))


def cpu_cg_baseline(w, threads, y, tol=1e-4, maxiter=300):
    """Oracle CG-SENSE solve (Toeplitz, lambda) to relative residual tol on
    the host: PSF + A^H y + iterations, repeated for about 10 s."""
    import oracle as O
    O.lib().orc_set_num_threads(threads)
    C = w["C"]
    M = w["R"] * w["S"]
    yh = y.cpu().numpy().reshape(C, M)
    reps, its, t0 = 0, 0, time.perf_counter()
    while True:
        ref = O.Nufft(w["img"], w["traj"], C, toeplitz=True)
        b = ref.sense_adjoint(w["maps"], yh)
        _, it, _, _ = ref.cgsense(w["maps"], b, w["lam"], maxiter, tol)
        its = it
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= 10.0 or reps >= 50:
            break
    X, Y, Z = w["img"]
    return dict(value=round(reps / dt, 4), unit="solves/s", cores=threads, kind="port", iterations=its,
                sample=f"oracle (BART-algorithm CPU restatement, not BART) PSF + A^H y + Toeplitz CG to relative "
                       f"residual {tol:g} ({its} iterations), {C} coils, {X}x{Y}x{Z}, {reps} solves ({dt:.1f} s)")


def cpu_baseline(w, threads):
    """Oracle NUFFT adj+fwd on a bounded sample: 8 coils of the same
    trajectory and image grid, repeated until about 10 s of CPU time."""
    import oracle as O
    O.lib().orc_set_num_threads(threads)
    C = min(8, w["C"])
    S = w["S"]
    M = w["R"] * S
    traj = w["traj"][:M]
    nf = O.Nufft(w["img"], traj, C, toeplitz=False)
    rng = np.random.default_rng(0)
    X, Y, Z = w["img"]
    ksp = (rng.standard_normal((C, M)) + 1j * rng.standard_normal((C, M))).astype(np.complex64)
    reps = 0
    t0 = time.perf_counter()
    while True:
        cim = nf.adjoint(ksp)
        nf.forward(cim)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= 10.0 or reps >= 200:
            break
    return dict(value=2 * C * M * reps / dt / 1e6, unit="Msamples/s", cores=threads, kind="port",
                sample=f"oracle adjoint+forward NUFFT, {C} coils x {S} spokes x {w['R']} samples on the "
                       f"{X}x{Y}x{Z} grid, {reps} repetitions ({dt:.1f} s)")


def run_batched(args, L, report):
    """96 independent CG-SENSE solves (30 iterations each, lambda 0.01): the
    right-hand sides A^H y, then one lsqr_conjgrad_batch call.  value = solves/s."""
    import torch
    import bart_amd as ba
    from bart_amd import inputs as I

    dev = torch.device("cuda:0")
    sl = I.batched_slices()
    iters, lam = 30, 0.01
    t0 = time.perf_counter()
    ops, ys, bs = [], [], []
    for d in sl:
        n, C, R, S = d["n"], d["C"], d["R"], d["S"]
        traj = I.radial_traj(R, S)
        img = I.ellipse_phantom(n, seed=d["seed"])
        maps = I.gaussian_coils((1, n, n), C, seed=d["seed"])
        op = ba.sense_nc_init(dims16(n, n, 1, C), dims16(n, n, 1, C), maps, dims16(1, R, S, C), dims16(3, R, S),
                              ba.traj_array(traj))
        x = torch.from_numpy(img.reshape(-1)).to(dev)
        y = op.forward(x)
        scale = float(y.abs().max())
        g = torch.Generator(device="cpu").manual_seed(d["seed"])
        noise = (torch.randn(y.numel(), generator=g) + 1j * torch.randn(y.numel(), generator=g)).to(torch.complex64)
        ys.append((y + (1e-3 * scale / np.sqrt(2.0)) * noise.to(dev)).contiguous())
        bs.append(torch.empty(n * n, dtype=torch.complex64, device=dev))
        ops.append(op)
    torch.cuda.synchronize()
    print(f"# {len(sl)} slices: plans and data in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    def step(rec):
        t0 = time.perf_counter()
        for op, y, b in zip(ops, ys, bs):
            op.adjoint(y, out=b)
        t1 = time.perf_counter()
        xs, st, tr = ba.lsqr_conjgrad_batch(ops, bs, iters, lam, 0.0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        rec.append((t1 - t0, t2 - t1, max(s.iterations for s in st),
                    max(s.resid / max(s.bnorm, 1e-30) for s in st)))

    for _ in range(args.warmup):
        step([])
    rec = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(rec)
        print(f"# step {i}: A^H y {rec[-1][0] * 1e3:.1f} ms, batched CG {rec[-1][1] * 1e3:.1f} ms "
              f"({rec[-1][2]} it, max rel-res {rec[-1][3]:.2e})", file=sys.stderr, flush=True)
    t_total = time.perf_counter() - t_start
    t_adj = sum(r[0] for r in rec) / len(rec)
    t_cg = sum(r[1] for r in rec) / len(rec)

    # the same solves one after another (lsqr_conjgrad per slice)
    t0 = time.perf_counter()
    for op, b in zip(ops, bs):
        ba.lsqr_conjgrad(op, b, iters, lam, 0.0)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0

    # compulsory HBM bytes of the solves: A^H y (k-space, coordinates, maps
    # in, image out) + per iteration image in/out, maps, PSF and the CG
    # vector updates (8 N (9 + C + 4) bytes for a 2-D slice)
    byts = 0
    for d in sl:
        N, C, M = d["n"] ** 2, d["C"], d["R"] * d["S"]
        byts += 8 * C * M + 8 * M + 8 * N + 8 * C * N + iters * 8 * N * (9 + C + 4)
    ach = byts / (t_adj + t_cg) / 1e9
    roofline = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, kernel="whole batched solve (end to end)",
                    bytes_per_launch=byts)

    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.lib().orc_set_num_threads(threads)
        done, t0 = 0, time.perf_counter()
        for d in sl:
            n, C, R, S = d["n"], d["C"], d["R"], d["S"]
            traj = I.radial_traj(R, S)
            maps = I.gaussian_coils((1, n, n), C, seed=d["seed"])
            ref = O.Nufft((n, n, 1), traj, C, toeplitz=True)
            b = ref.sense_adjoint(maps, ys[done].cpu().numpy())
            ref.cgsense(maps, b, lam, iters, 0.0)
            done += 1
            if time.perf_counter() - t0 >= 10.0:
                break
        dt = time.perf_counter() - t0
        cpu = dict(value=round(done / dt, 4), unit="solves/s", cores=threads, kind="port",
                   sample=f"oracle PSF + A^H y + {iters} Toeplitz CG iterations for the first {done} slices of "
                          f"the batch ({dt:.1f} s, plan creation included)")

    tot_samples = sum(d["C"] * d["R"] * d["S"] for d in sl)
    out = {
        "metric": "NUFFT adj+fwd Msamples/s; CG-SENSE solves/s to rel-res 1e-4, warm vs cold",
        "value": round(len(sl) / (t_adj + t_cg), 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex; fp64 CG reductions)",
        "data": "synthetic",
        "config": {"workload": "batched_uneven_2d_radial_96_slices", "slices": len(sl),
                   "matrix": sorted(set(d["n"] for d in sl)), "coils": sorted(set(d["C"] for d in sl)),
                   "coil_samples": tot_samples, "cg_iters": iters, "lambda": lam},
        "adjoint_rhs_ms": round(t_adj * 1e3, 3),
        "batched_cg_ms": round(t_cg * 1e3, 3),
        "sequential_cg_ms": round(t_seq * 1e3, 3),
        "cg_iterations": rec[-1][2],
        "grouped_launches": os.environ.get("BART_AMD_BATCH_GROUPS", "on") != "off",
        "autotuned_adjoint_gridding": ba.tuning_report(),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_warmstart(args):
    """Learned warm start (2-D radial, 256^2, 8 coils, 32-96 spokes): training
    pairs from this solver (CG to relative residual 1e-6), training of the
    6-layer network, then CG-SENSE to relative residual 1e-4 on held-out
    seeds from zero and from the network's guess.  value = warm solves/s."""
    import torch
    from bart_amd import warmstart as WS

    n, lam = 256, 0.01
    train_seeds = list(range(args.ws_train))
    eval_seeds = list(range(100000, 100000 + args.ws_eval))
    t0 = time.perf_counter()
    inp, tgt, its = WS.training_pairs(train_seeds, n=n, lam=lam,
                                      log=lambda m: print(f"# {m}", file=sys.stderr, flush=True))
    t_data = time.perf_counter() - t0
    print(f"# {len(train_seeds)} training pairs in {t_data:.1f} s (CG iterations to 1e-6: median "
          f"{int(np.median(its))}, max {max(its)})", file=sys.stderr, flush=True)
    net = WS.WarmNet(48, seed=0)
    t0 = time.perf_counter()
    hist = WS.train(net, inp, tgt, epochs=args.ws_epochs, batch=8, lr=1e-3,
                    log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    hist_r = []
    if args.ws_finetune > 0:
        # fine-tuning on the normal-equation residual of the guess
        probs = WS.residual_problems(train_seeds[:args.ws_finetune], n=n, lam=lam)
        hist_r = WS.train_residual(net, probs, n, epochs=args.ws_ft_epochs, lr=2e-4, lam=lam,
                                   log=lambda m: print("# " + m, file=sys.stderr, flush=True))
        for p in probs:
            p[0].free()
        del probs
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0

    # network forward on one 256^2 image: 2 9 (2 48 + 4 48 48 + 48 2) n^2 flops
    flops = 2 * 9 * (2 * 48 + 4 * 48 * 48 + 48 * 2) * n * n
    xb = inp[:8].contiguous()
    net.forward(xb)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        net.forward(xb)
    e1.record()
    torch.cuda.synchronize()
    t_fwd = e0.elapsed_time(e1) / 5 / 8
    del inp, tgt

    res = []
    for s in eval_seeds:
        op, b, _ = WS.make_problem(s, n=n, lam=lam)
        res.append(WS.solve_warm_and_cold(net, op, b, n, lam=lam))
        op.free()
        if len(res) % 64 == 0:
            print(f"# eval: {len(res)} problems", file=sys.stderr, flush=True)
    cold_ms = sum(r["cold_ms"] for r in res)
    warm_ms = sum(r["warm_ms"] for r in res)
    print(f"# eval: cold {np.mean([r['cold_it'] for r in res]):.1f} it, warm {np.mean([r['warm_it'] for r in res]):.1f} it, "
          f"fallbacks {sum(r['fell_back'] for r in res)}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        import oracle as O
        from bart_amd import inputs as I
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.lib().orc_set_num_threads(threads)
        done, t0 = 0, time.perf_counter()
        for s in eval_seeds:
            op, b, _ = WS.make_problem(s, n=n, lam=lam)
            rng = np.random.default_rng(s)
            S = int(rng.integers(32, 97))
            maps = I.gaussian_coils((1, n, n), 8, seed=s)
            ref = O.Nufft((n, n, 1), I.radial_traj(2 * n, S), 8, toeplitz=True)
            ref.cgsense(maps, b.cpu().numpy().reshape(1, n, n), lam, 300, 1e-4)
            op.free()
            done += 1
            if time.perf_counter() - t0 >= 10.0:
                break
        dt = time.perf_counter() - t0
        cpu = dict(value=round(done / dt, 4), unit="solves/s", cores=threads, kind="port",
                   sample=f"oracle PSF + Toeplitz CG-SENSE to rel-res 1e-4 from zero on the first {done} held-out "
                          f"problems ({dt:.1f} s)")

    ach = flops / (t_fwd * 1e-3) / 1e12
    out = {
        "metric": "NUFFT adj+fwd Msamples/s; CG-SENSE solves/s to rel-res 1e-4, warm vs cold",
        "value": round(len(res) / (warm_ms * 1e-3), 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(res),
        "warmup": 0,
        "ms_per_step": round(warm_ms / len(res), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 CG (fp64 reductions); bf16 network with fp32 accumulation",
        "data": "synthetic",
        "config": {"workload": "learned_warm_start_2d_radial_256^2_8coils_32-96_spokes", "train_pairs": len(train_seeds),
                   "eval_problems": len(res), "epochs": args.ws_epochs, "network": "6 x conv3x3, 48 channels, residual",
                   "tolerance": 1e-4, "lambda": lam},
        "cold_solves_per_s": round(len(res) / (cold_ms * 1e-3), 3),
        "time_to_tolerance_cold_over_warm": round(cold_ms / warm_ms, 3),
        "iterations_cold_mean": float(np.mean([r["cold_it"] for r in res])),
        "iterations_warm_mean": float(np.mean([r["warm_it"] for r in res])),
        "fallbacks": int(sum(r["fell_back"] for r in res)),
        "warm_vs_cold_solution_relerr_max": float(max(r["warm_vs_cold"] for r in res)),
        "guess_ms_mean": round(float(np.mean([r["guess_ms"] for r in res])), 3),
        "training_data_s": round(t_data, 2),
        "training_s": round(t_train, 2),
        "train_loss_first_last": [hist[0], hist[-1]],
        "residual_finetune": {"problems": args.ws_finetune, "epochs": args.ws_ft_epochs,
                              "loss_first_last": [hist_r[0], hist_r[-1]] if hist_r else None},
        "roofline": {"bound": "mfma", "achieved": round(ach, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(ach / 2500.0, 5), "traffic": None, "kernel": "network forward (6 conv3x3 layers)",
                     "flops_per_image": flops, "ms_per_image": round(t_fwd, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="kooshball", choices=["kooshball", "radial2d", "spiral", "batched", "warmstart", "widened"])
    ap.add_argument("--ws-train", type=int, default=1024, help="warmstart: training pairs")
    ap.add_argument("--ws-eval", type=int, default=32, help="warmstart: held-out problems")
    ap.add_argument("--ws-epochs", type=int, default=40, help="warmstart: training epochs")
    ap.add_argument("--ws-finetune", type=int, default=0,
                    help="warmstart: problems for fine-tuning on the normal-equation residual (0: none)")
    ap.add_argument("--ws-ft-epochs", type=int, default=10, help="warmstart: residual fine-tuning epochs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-radial", action="store_true",
                    help="kooshball: skip the 2-D radial CG-SENSE solves to tolerance")
    ap.add_argument("--tol-solves", type=int, default=3, help="CG-SENSE solves to 1e-4 timed (median)")
    ap.add_argument("--tol-maxiter", type=int, default=200, help="iteration cap of the solves to 1e-4")
    ap.add_argument("--radial-solves", type=int, default=10, help="2-D radial solves to 1e-4 timed (median)")
    ap.add_argument("--radial-warm-pairs", type=int, default=256,
                    help="2-D radial warm start: training pairs of the same protocol (0: no warm start)")
    ap.add_argument("--radial-warm-epochs", type=int, default=10, help="2-D radial warm start: training epochs")
    ap.add_argument("--no-profile", action="store_true", help="do not time kernels with HIP events")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="hardware-counter target: run the adjoint, forward and SENSE normal operator twice "
                         "and a 1 GiB device copy (the counters' calibration), then exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {args.gpus} is not supported "
              "(sharding the problems or coil groups over GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    import torch
    import ctypes
    import bart_amd as ba

    assert torch.cuda.is_available(), "bench needs a GPU"
    dev = torch.device("cuda:0")
    L = ba.lib()
    L.bart_amd_profile.argtypes = [ctypes.c_int]
    L.bart_amd_profile_report.restype = ctypes.c_int
    L.bart_amd_profile_report.argtypes = [ctypes.c_char_p, ctypes.c_int]

    if args.config == "batched":
        run_batched(args, L, None)
        return

    if args.config == "warmstart":
        run_warmstart(args)
        return

    if args.config == "widened":
        run_widened(args, L, torch.device("cuda:0"))
        return

    t_gen = time.perf_counter()
    w = workload(args.config)
    X, Y, Z = w["img"]
    C, R, S = w["C"], w["R"], w["S"]
    M = R * S
    N = X * Y * Z
    print(f"# workload {w['name']}: generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)

    traj = ba.traj_array(w["traj"])
    conf = ba.NufftConf.defaults(toeplitz=False)
    t0 = time.perf_counter()
    nufft = ba.nufft_create2(dims16(1, R, S, C), dims16(X, Y, Z, C), dims16(3, R, S), traj, conf=conf)
    sense = ba.sense_nc_init(dims16(X, Y, Z, C), dims16(X, Y, Z, C), w["maps"], dims16(1, R, S, C),
                             dims16(3, R, S), traj)
    t_plan = time.perf_counter() - t0
    print(f"# plans (binning + PSF) in {t_plan:.2f} s", file=sys.stderr, flush=True)

    x = torch.from_numpy(w["x"].reshape(-1)).to(dev)

    if args.pmc_pass:
        yk = sense.forward(x)
        ci = torch.empty(C * N, dtype=torch.complex64, device=dev)
        k2 = torch.empty(C * M, dtype=torch.complex64, device=dev)
        o = torch.empty(N, dtype=torch.complex64, device=dev)
        for _ in range(2):
            nufft.adjoint(yk, out=ci)
            nufft.forward(ci, out=k2)
            sense.normal(x, out=o)
        torch.cuda.synchronize()
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        a2 = torch.empty_like(a)
        for _ in range(2):
            a2.copy_(a)
        torch.cuda.synchronize()
        print("# pmc pass done", file=sys.stderr, flush=True)
        return

    y = sense.forward(x)
    scale = float(y.abs().max())
    g = torch.Generator(device="cpu").manual_seed(0)
    noise = (torch.randn(M * C, generator=g) + 1j * torch.randn(M * C, generator=g)).to(torch.complex64)
    y = (y + (1e-3 * scale / np.sqrt(2.0)) * noise.to(dev)).contiguous()
    cim = torch.empty(C * N, dtype=torch.complex64, device=dev)
    ksp2 = torch.empty(C * M, dtype=torch.complex64, device=dev)
    b = torch.empty(N, dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()

    def step(rec):
        t0 = time.perf_counter()
        nufft.adjoint(y, out=cim)
        nufft.forward(cim, out=ksp2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sense.adjoint(y, out=b)
        xs, st, tr = ba.lsqr_conjgrad(sense, b, w["iters"], w["lam"], 0.0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        rec.append((t1 - t0, t2 - t1, st.iterations, st.resid / max(st.bnorm, 1e-30)))

    warm = []
    for _ in range(args.warmup):
        step(warm)
    torch.cuda.synchronize()

    if not args.no_profile:
        L.bart_amd_profile(1)
    rec = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(rec)
        print(f"# step {i}: nufft adj+fwd {rec[-1][0] * 1e3:.2f} ms, cg-sense {rec[-1][1] * 1e3:.2f} ms "
              f"({rec[-1][2]} it, rel-res {rec[-1][3]:.2e})", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    L.bart_amd_profile(0)

    def report():
        n = L.bart_amd_profile_report(None, 0)
        buf = ctypes.create_string_buffer(n + 16)
        L.bart_amd_profile_report(buf, n + 16)
        return json.loads(buf.value.decode())

    prof = {}
    prof_n = {}
    n_norm = 5
    if not args.no_profile:
        prof = report()
        # the CG iterations run as one HIP graph, whose kernels the event
        # profiler does not see: time the same SENSE normal operator outside
        # the graph and count it once per CG iteration of the step
        tmp = torch.empty(N, dtype=torch.complex64, device=dev)
        sense.normal(b, out=tmp)
        torch.cuda.synchronize()
        L.bart_amd_profile(1)
        for _ in range(n_norm):
            sense.normal(b, out=tmp)
        torch.cuda.synchronize()
        L.bart_amd_profile(0)
        prof_n = report()

    t_nufft = sum(r[0] for r in rec) / len(rec)
    t_cg = sum(r[1] for r in rec) / len(rec)
    its = rec[-1][2]
    relres = rec[-1][3]
    value = 2 * C * M / t_nufft / 1e6

    # CG-SENSE to relative residual 1e-4 (the metric's solve), cold start
    tol_rec = []
    for _ in range(args.tol_solves):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sense.adjoint(y, out=b)
        xs, st, tr = ba.lsqr_conjgrad(sense, b, args.tol_maxiter, w["lam"], 1e-4)
        torch.cuda.synchronize()
        tol_rec.append((time.perf_counter() - t0, st.iterations, st.status, st.resid / max(st.bnorm, 1e-30)))
    t_tol = float(np.median([r[0] for r in tol_rec]))
    print(f"# cg-sense to 1e-4: {tol_rec[-1][1]} it, {t_tol * 1e3:.2f} ms", file=sys.stderr, flush=True)

    # per step: (ms, launches, mean ms per launch) of every kernel class
    per_step = {}
    for k, (cnt, tot, mn, mx) in prof.items():
        per_step[k] = [tot / args.steps, cnt / args.steps, tot / cnt]
    for k, (cnt, tot, mn, mx) in prof_n.items():
        launches = cnt / n_norm * its
        e = per_step.setdefault(k, [0.0, 0.0, tot / cnt])
        e[0] += tot / cnt * launches
        e[1] += launches

    sb = survey_bytes(w)
    roofline = None
    if per_step:
        name, (ms_step, launches, avg_ms) = max(per_step.items(), key=lambda kv: kv[1][0])
        unit = UNIT_OF.get(name)
        kb = algorithmic_bytes(name, w)
        if unit is not None:
            bts = sb[unit]
            ach = bts / (avg_ms * 1e-3) / 1e9
            tr = TRAFFIC.get((w["name"], name))
            roofline = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                            frac=round(ach / HBM_PEAK_GBS, 4), traffic=tr, kernel=name,
                            bytes_per_launch=bts, bytes_unit=f"one {unit} (compulsory bytes)",
                            avg_ms=round(avg_ms, 4), launches_per_step=round(launches, 1),
                            share_of_step=round(ms_step / (t_total / args.steps * 1e3), 3),
                            kernel_bytes_per_launch=kb,
                            kernel_bytes_frac=(round(kb / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if kb else None),
                            traffic_source="rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py --pmc-pass "
                                           "(profiles/)" if tr else None)

    # the whole Toeplitz CG iteration and the NUFFT pair against their
    # compulsory bytes (end to end, all kernels)
    iter_ms = sum(v[0] for k, v in per_step.items() if UNIT_OF.get(k) == "iteration" or k.startswith("cg_")) / max(its, 1)
    e2e = dict(nufft_adj_fwd=dict(ms=round(t_nufft * 1e3, 3), bytes=2 * sb["nufft"],
                                  frac=round(2 * sb["nufft"] / t_nufft / 1e9 / HBM_PEAK_GBS, 4)),
               cg_iteration=(dict(ms=round(iter_ms, 4), bytes=sb["iteration"],
                                  frac=round(sb["iteration"] / (iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
                             if iter_ms > 0 else None),
               solve_to_tol=dict(ms=round(t_tol * 1e3, 3), bytes=sb["sense_adjoint"] + tol_rec[-1][1] * sb["iteration"],
                                 frac=round((sb["sense_adjoint"] + tol_rec[-1][1] * sb["iteration"]) / t_tol / 1e9
                                            / HBM_PEAK_GBS, 4)))

    # accuracy: FP64 direct NUDFT on a sample subset (8-coil oracle leg aside)
    accuracy = None
    if not args.no_cpu:
        accuracy = fp64_nudft_check(w, cim, ksp2, 64 if Z > 1 else 2048)
        print(f"# fp64 NUDFT check: {accuracy}", file=sys.stderr, flush=True)

    # explicit grid/gridH normal operator against the Toeplitz one (2-D)
    normal_cmp = None
    if Z == 1 and not args.no_cpu:
        normal_cmp = explicit_vs_toeplitz(w, b, sense)

    # the 2-D radial CG-SENSE solve of the metric (256^2, 8 coils, 201 x 512)
    # to relative residual 1e-4, cold and from the learned warm start
    radial = None
    if args.config == "kooshball" and not args.no_radial:
        radial = radial_solves(args, dev)

    cpu = None
    cpu_cg = None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if not args.no_cpu:
        cpu = cpu_baseline(w, threads)
        cpu["probe"] = cpu_probe()
        if Z == 1:
            cpu_cg = cpu_cg_baseline(w, threads, y)
        elif radial is not None:
            cpu_cg = radial.pop("cpu", None)

    out = {
        "metric": "NUFFT adj+fwd Msamples/s; CG-SENSE solves/s to rel-res 1e-4, warm vs cold",
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 complex; fp64 CG reductions)",
        "data": "synthetic",
        "config": {"workload": w["name"], "coils": C, "samples_per_coil": M, "image": list(w["img"]),
                   "oversampling": 2.0, "kb_width": 6, "cg_iters": w["iters"], "lambda": w["lam"]},
        "nufft_adj_fwd_ms": round(t_nufft * 1e3, 3),
        "cg_sense_fixed_iters": dict(iterations=its, ms=round(t_cg * 1e3, 3), solves_per_s=round(1.0 / t_cg, 3),
                                     rel_res=relres),
        "cg_sense_to_tol_cold": dict(tol=1e-4, iterations=tol_rec[-1][1], converged=tol_rec[-1][2] == 2,
                                     rel_res=tol_rec[-1][3], ms=round(t_tol * 1e3, 3),
                                     solves_per_s=round(1.0 / t_tol, 3)),
        "radial2d_cg_sense_to_tol": radial,
        "roofline": roofline,
        "end_to_end_vs_compulsory_bytes": e2e,
        "cpu_baseline": cpu,
        "cpu_baseline_cg_sense": cpu_cg,
        "fp64_nudft": accuracy,
        "normal_explicit_vs_toeplitz": normal_cmp,
        "kernels_ms_per_step": {k: round(v[0], 4) for k, v in sorted(per_step.items())},
        "autotuned_adjoint_gridding": ba.tuning_report(),
    }
    if cpu_cg and radial and radial.get("cold"):
        out["cg_sense_gpu_over_cpu_radial2d"] = round(radial["cold"]["solves_per_s"] / cpu_cg["value"], 1)
    elif cpu_cg and Z == 1:
        out["cg_sense_gpu_over_cpu"] = round(1.0 / t_tol / cpu_cg["value"], 1)
    print(json.dumps(out), flush=True)


def _median_ms(fn, reps):
    """Median wall time of fn() in ms over reps calls after one untimed
    call, each bracketed by device synchronisations."""
    import torch
    fn()
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)) * 1e3


def _cpu_ms(fn, budget_s=4.0, max_reps=50):
    """Mean wall time of fn() in ms, repeated for about budget_s."""
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or reps >= max_reps:
            return dt / reps * 1e3, reps


def _frac(nbytes, ms):
    return round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)


def run_widened(args, L, dev):
    """The rows beyond the plain NUFFT / CG-SENSE path, each on a named
    workload, timed on the GPU (median wall time with the device
    synchronised, inputs resident) beside the oracle on the host cores:
      * subspace basis: 2-D radial 256^2, 8 coils, 64 frames of 4 spokes x 512
        samples, K = 4 coefficients: NUFFT adjoint + forward, the Toeplitz
        normal operator with the K x K PSF, 30 CG-SENSE iterations;
      * chained NUFFT (os 1.5, width 6) on the 2-D radial 256^2 / 8 coils /
        201 x 512 workload: adjoint + forward;
      * Strang preconditioner: the same workload's inverse NUFFT (nufft -i),
        10 and 30 iterations with and without it, and the residual of the
        normal equations each leaves;
      * ESPIRiT: 2-D 256^2 / 8 coils and 3-D 128^3 / 16 coils, 2 maps, 24^3
        calibration region, 6^3 kernel (BART's defaults), maps and
        eigenvalue maps into device memory.
    Each row carries its compulsory bytes (inputs read once, outputs written
    once) and the fraction of the HBM peak that its time corresponds to.
    Prints one JSON line."""
    import ctypes
    import torch
    import bart_amd as ba
    from bart_amd import inputs as I

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cpu = not args.no_cpu
    if cpu:
        import oracle as O
        O.lib().orc_set_num_threads(threads)
    rows = {}

    def crandn(shape, seed):
        r = np.random.default_rng(seed)
        return (r.standard_normal(shape) + 1j * r.standard_normal(shape)).astype(np.complex64)

    def on_dev(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.complex64).ravel()).to(dev)

    # -- subspace basis --------------------------------------------------
    n, C, T, spf, R, K = 256, 8, 64, 4, 512, 4
    N = n * n
    Mf = spf * R
    traj = I.radial_traj(R, T * spf).reshape(T, Mf, 3)
    rng = np.random.default_rng(11)
    B = np.linalg.qr(rng.standard_normal((T, K)) + 1j * rng.standard_normal((T, K)))[0].astype(np.complex64)
    conf = ba.NufftConf.defaults()
    bas = np.ascontiguousarray(B.T)
    t0 = time.perf_counter()
    op = ba.nufft_create2(dims16(1, Mf, 1, C, 1, T), dims16(n, n, 1, C, 1, 1, K), dims16(3, Mf, 1, 1, 1, T),
                          np.ascontiguousarray(traj), conf=conf, bas_dims=dims16(1, 1, 1, 1, 1, T, K), basis=bas)
    t_plan = time.perf_counter() - t0
    x = on_dev(crandn((K, C, 1, n, n), 12))
    y = on_dev(crandn((T, C, Mf), 13))
    xo = torch.empty_like(x)
    yo = torch.empty_like(y)
    t_fa = _median_ms(lambda: (op.adjoint(y, out=xo), op.forward(xo, out=yo)), 10)
    t_n = _median_ms(lambda: op.normal(x, out=xo), 20)
    maps = I.gaussian_coils((1, n, n), C, seed=14)
    sense = ba.sense_nc_init(dims16(n, n, 1, C, 1, 1, K), dims16(n, n, 1, C), maps, dims16(1, Mf, 1, C, 1, T),
                             dims16(3, Mf, 1, 1, 1, T), np.ascontiguousarray(traj), basis_dims=dims16(1, 1, 1, 1, 1, T, K),
                             basis=bas)
    b = sense.adjoint(y)
    t_cg = _median_ms(lambda: ba.lsqr_conjgrad(sense, b, 30, 0.01, 0.0), 5)
    nb_fa = 2 * (8 * C * T * Mf + 4 * 2 * T * Mf + 8 * C * K * N)
    nb_n = 2 * 8 * C * K * N + 8 * K * K * 4 * N
    row = dict(workload=f"2d_radial_subspace_{n}^2_{C}coils_{T}frames_{spf}x{R}_K{K}", plan_s=round(t_plan, 2),
               adj_fwd_ms=round(t_fa, 3), adj_fwd_msamples_s=round(2 * C * T * Mf / t_fa / 1e3, 1),
               adj_fwd_bytes=nb_fa, adj_fwd_frac=_frac(nb_fa, t_fa),
               normal_ms=round(t_n, 3), normal_bytes=nb_n, normal_frac=_frac(nb_n, t_n),
               cg_sense_30_ms=round(t_cg, 3))
    if cpu:
        ref = O.BasisNufft((n, n, 1), traj, B, C, toeplitz=True)
        xh = crandn((K, C, 1, n, n), 12)
        yh = crandn((T, C, Mf), 13)
        c_fa, r1 = _cpu_ms(lambda: ref.forward(ref.adjoint(yh)))
        c_n, r2 = _cpu_ms(lambda: ref.normal(xh))
        row["cpu"] = dict(adj_fwd_ms=round(c_fa, 1), normal_ms=round(c_n, 1), cores=threads, kind="port",
                          sample=f"oracle BasisNufft, same workload ({r1} / {r2} repetitions)")
    rows["basis"] = row
    print(f"# basis: {row}", file=sys.stderr, flush=True)
    sense.free()
    op.free()

    # -- chained NUFFT and the preconditioner (2-D radial) -----------------
    w = workload("radial2d")
    X, Y, Z = w["img"]
    C, R, S = w["C"], w["R"], w["S"]
    M, N = R * S, X * Y * Z
    tr = ba.traj_array(w["traj"])
    conf = ba.NufftConf.defaults(toeplitz=False)
    conf.os = 1.5
    op = ba.nufft_create2(dims16(1, R, S, C), dims16(X, Y, Z, C), dims16(3, R, S), tr, conf=conf)
    y = on_dev(crandn((C, M), 21))
    ci = torch.empty(C * N, dtype=torch.complex64, device=dev)
    ko = torch.empty_like(y)
    t_fa = _median_ms(lambda: (op.adjoint(y, out=ci), op.forward(ci, out=ko)), 20)
    nb = 2 * (8 * C * M + 4 * 2 * M + 8 * C * N)
    row = dict(workload=w["name"] + "_os1.5_width6", adj_fwd_ms=round(t_fa, 3),
               adj_fwd_msamples_s=round(2 * C * M / t_fa / 1e3, 1), adj_fwd_bytes=nb, adj_fwd_frac=_frac(nb, t_fa))
    if cpu:
        ref = O.ChainNufft(w["img"], w["traj"], C, os_=1.5, width=6.0)
        yh = crandn((C, M), 21)
        c_fa, r1 = _cpu_ms(lambda: ref.forward(ref.adjoint(yh)))
        row["cpu"] = dict(adj_fwd_ms=round(c_fa, 1), cores=threads, kind="port",
                          sample=f"oracle ChainNufft, same workload ({r1} repetitions)")
    rows["chain"] = row
    print(f"# chain: {row}", file=sys.stderr, flush=True)
    op.free()

    conf = ba.NufftConf.defaults()
    op = ba.nufft_create2(dims16(1, R, S, C), dims16(X, Y, Z, C), dims16(3, R, S), tr, conf=conf)
    maps = np.asarray(w["maps"]).reshape(C, N)
    yk = op.forward(on_dev(maps * w["x"].reshape(1, N)))
    bh = op.adjoint(yk).cpu().numpy()
    pre = ba.Precond(op)
    bd = on_dev(bh)
    res = {}
    for its in (10, 30):
        for name, fn in (("plain", lambda: ba.lsqr_conjgrad(op, bh, its, 0.01, 0.0)),
                         ("strang", lambda: ba.lsqr_conjgrad_precond(op, pre, bh, its, 0.01, 0.0))):
            t = _median_ms(fn, 3)
            xs = on_dev(np.asarray(fn()[0]))
            r = bd - op.normal(xs) - 0.01 * xs
            res[f"{name}_{its}"] = dict(ms=round(t, 3), rel_res=float(torch.linalg.norm(r) / torch.linalg.norm(bd)))
    # CG on P (A^H A + lambda I), as lsqr chains the preconditioner
    # (iter/lsqr.c:126-142); rel_res is || b - (A^H A + lambda I) x || / || b ||
    rows["precond"] = dict(workload=w["name"] + "_inverse_nufft_lambda0.01", **res)
    print(f"# precond: {rows['precond']}", file=sys.stderr, flush=True)
    pre.free()
    op.free()

    # -- ESPIRiT -----------------------------------------------------------
    L.extract_calib.restype = ctypes.c_void_p
    L.extract_calib.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_bool]
    L.calib.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]

    def cart_ksp(dims, C, seed):
        X, Y, Z = dims
        img = (I.ellipse_phantom(X, seed=seed) if Z == 1 else I.ellipsoid_phantom(X, nell=8, seed=seed))
        Sm = I.gaussian_coils((Z, Y, X), C, seed=seed + 1)
        ax = (1, 2, 3) if Z > 1 else (2, 3)
        k = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(Sm * img[None], axes=ax), axes=ax, norm="ortho"), axes=ax)
        return k.astype(np.complex64)

    for dims, C in (((256, 256, 1), 8), ((128, 128, 128), 16)):
        X, Y, Z = dims
        Mp = 2
        ksp = cart_ksp(dims, C, 31)
        kd = (ctypes.c_long * 16)(*dims16(X, Y, Z, C))
        cd = (ctypes.c_long * 16)()
        cs = (ctypes.c_long * 3)(24, 24, 24)
        ec = ba.EcalibConf.defaults(kdims=[6 if v > 1 else 1 for v in dims])
        od = (ctypes.c_long * 16)(*dims16(X, Y, Z, C, Mp))
        Np = X * Y * Z
        mo = torch.empty(Mp * C * Np, dtype=torch.complex64, device=dev)
        eo = torch.empty(Mp * Np, dtype=torch.complex64, device=dev)
        nsv = 216 * C

        def espirit():
            p = L.extract_calib(cd, cs, kd, ksp.ctypes.data, False)
            L.calib(ctypes.byref(ec), od, mo.data_ptr(), eo.data_ptr(), 0, None, cd, p)
            ctypes.CDLL(None).free(ctypes.c_void_p(p))

        t_e = _median_ms(espirit, 3)
        ba._check()
        nb = 8 * C * Np + 8 * Mp * C * Np + 8 * Mp * Np
        row = dict(workload=f"espirit_{X}x{Y}x{Z}_{C}coils_{Mp}maps_cal24_kernel6", ms=round(t_e, 2),
                   bytes=nb, frac=_frac(nb, t_e))
        if cpu and Z == 1:
            from oracle import espirit as E
            c_e, r1 = _cpu_ms(lambda: E.ecalib(ksp, maps=Mp), budget_s=8.0, max_reps=5)
            row["cpu"] = dict(ms=round(c_e, 1), cores=threads, kind="port",
                              sample=f"oracle ecalib (float64 numpy), same workload ({r1} repetitions)")
        rows["espirit_" + ("2d" if Z == 1 else "3d")] = row
        print(f"# espirit: {row}", file=sys.stderr, flush=True)

    print(json.dumps(dict(metric="rows beyond the NUFFT / CG-SENSE path (basis, chain, preconditioner, ESPIRiT)",
                          n_gpus=1, data="synthetic", rows=rows)))


def radial_solves(args, dev):
    """CG-SENSE on the 2-D radial workload (256^2, 8 coils, 201 x 512,
    lambda 0.01) to relative residual 1e-4: A^H y then CG from zero, and
    from the learned warm start (network trained in this run, unless
    --radial-warm-pairs 0).  Also the oracle's CPU solve of the same problem
    (about 10 s)."""
    import torch
    import bart_amd as ba
    w = workload("radial2d")
    X, Y, Z = w["img"]
    C, R, S = w["C"], w["R"], w["S"]
    M = R * S
    N = X * Y * Z
    sense = ba.sense_nc_init(dims16(X, Y, Z, C), dims16(X, Y, Z, C), w["maps"], dims16(1, R, S, C),
                             dims16(3, R, S), ba.traj_array(w["traj"]))
    x = torch.from_numpy(w["x"].reshape(-1)).to(dev)
    y = sense.forward(x)
    scale = float(y.abs().max())
    g = torch.Generator(device="cpu").manual_seed(1)
    noise = (torch.randn(M * C, generator=g) + 1j * torch.randn(M * C, generator=g)).to(torch.complex64)
    y = (y + (1e-3 * scale / np.sqrt(2.0)) * noise.to(dev)).contiguous()
    b = torch.empty(N, dtype=torch.complex64, device=dev)

    def cold():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sense.adjoint(y, out=b)
        xs, st, _ = ba.lsqr_conjgrad(sense, b, 300, w["lam"], 1e-4)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, st

    cold()
    rec = [cold() for _ in range(args.radial_solves)]
    t = float(np.median([r[0] for r in rec]))
    st = rec[-1][1]
    out = dict(workload=w["name"], tol=1e-4,
               cold=dict(iterations=st.iterations, converged=st.status == 2, ms=round(t * 1e3, 3),
                         solves_per_s=round(1.0 / t, 2)),
               warm=None)
    print(f"# radial2d cg-sense to 1e-4 (cold): {st.iterations} it, {t * 1e3:.3f} ms", file=sys.stderr, flush=True)
    if args.radial_warm_pairs > 0:
        out["warm"] = radial_warm(args, sense, w, y, b, dev)
        print(f"# radial2d cg-sense to 1e-4 (warm): {out['warm']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu"] = cpu_cg_baseline(w, threads, y)
    sense.free()
    return out


def radial_warm(args, sense, w, y, b, dev):
    """The learned warm start for the 2-D radial solve: a network (6 layers,
    48 channels) trained in this run on pairs of the same protocol --
    the same trajectory, coils and lambda, other seeded ellipse phantoms with
    noise -- (input: the density-compensated adjoint A^H (w y), ramp w, at
    its best multiple, scaled; target: the CG solution to relative residual
    1e-6), then the bench's own problem solved to 1e-4 from the library's
    guess (bart_amd_warmnet_guess_dcf) with the zero-start fallback.
    Times include A^H y and the guess."""
    import torch
    import bart_amd as ba
    from bart_amd import inputs as I
    from bart_amd import warmstart as WS

    n = w["img"][0]
    C, lam = w["C"], w["lam"]
    M = w["R"] * w["S"]
    dcf = WS.radial_dcf(w["R"], w["S"])
    t0 = time.perf_counter()
    xin, xt = [], []
    for k in range(args.radial_warm_pairs):
        img = torch.from_numpy(I.ellipse_phantom(n, nell=10, seed=5000 + k).reshape(-1)).to(dev)
        yk = sense.forward(img)
        g = torch.Generator(device="cpu").manual_seed(5000 + k)
        e = (torch.randn(M * C, generator=g) + 1j * torch.randn(M * C, generator=g)).to(torch.complex64)
        yk = (yk + (1e-3 * float(yk.abs().max()) / np.sqrt(2.0)) * e.to(dev)).contiguous()
        bk = sense.adjoint(yk)
        inp, sc = WS.network_input(sense, bk, n, lam, WS.dcf_direction(sense, yk, dcf))
        xs, _, _ = ba.lsqr_conjgrad(sense, bk, 150, lam, 1e-6)
        xin.append(inp)
        xt.append(WS.to_channels(xs, n, sc))
    inputs, targets = torch.cat(xin), torch.cat(xt)
    t_data = time.perf_counter() - t0
    net = WS.WarmNet(48, seed=0)
    t0 = time.perf_counter()
    hist = WS.train(net, inputs, targets, epochs=args.radial_warm_epochs, batch=8, lr=1e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0

    def warm():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sense.adjoint(y, out=b)
        x0 = WS.guess_dcf(net, sense, b, lam, y, dcf)
        xs, st, _ = ba.lsqr_conjgrad(sense, b, 300, lam, 1e-4, x0=x0, fallback=True)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, st

    warm()
    rec = [warm() for _ in range(args.radial_solves)]
    t = float(np.median([r[0] for r in rec]))
    st = rec[-1][1]
    net.free()
    return dict(iterations=st.iterations, converged=st.status == 2, fell_back=bool(st.fell_back), ms=round(t * 1e3, 3),
                solves_per_s=round(1.0 / t, 2), training_pairs=args.radial_warm_pairs,
                training_epochs=args.radial_warm_epochs, train_loss_first_last=[hist[0], hist[-1]],
                training_data_s=round(t_data, 2), training_s=round(t_train, 2))


if __name__ == "__main__":
    main()
