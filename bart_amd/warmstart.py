"""Learned CG warm start: the network of the C ABI (bart_amd_warmnet_*) and
the host-side steps around it -- training pairs from this solver, training,
and CG from the network's guess with the zero start as fallback.

No counterpart in BART: pics takes an initial guess with -W
(src/pics.c:580-594, iter/itop.c:54-66); this module proposes it.  The
network's input is the density-compensated adjoint A^H (w y) (the
coil-combined gridding image, w the radial ramp) times its best multiple for
the residual of the normal equations; its output is the guess.  The network
runs on the device through the library (3x3 convolutions on the matrix
cores); torch only holds device memory.
"""
import ctypes
import time

import numpy as np

from . import BartError, _check, _fptr, lib

DIMS = 16


def _lib():
    return lib()


class WarmNet:
    """2 -> channels (x5) -> 2 channel residual CNN (ReLU, 3x3 convolutions)."""

    def __init__(self, channels=48, seed=0):
        self.h = _lib().bart_amd_warmnet_create(channels, seed)
        if not self.h:
            _check()
            raise BartError("warm-start network creation failed")
        self.channels = channels

    def free(self):
        if getattr(self, "h", None):
            _lib().bart_amd_warmnet_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def nparams(self):
        return int(_lib().bart_amd_warmnet_num_params(self.h))

    def _params(self, mode, values=None):
        buf = np.empty(self.nparams, np.float32) if values is None else np.ascontiguousarray(values, np.float32)
        _lib().bart_amd_warmnet_params(self.h, buf.ctypes.data_as(ctypes.c_void_p), mode)
        _check()
        return buf

    def get_params(self):
        return self._params(0)

    def set_params(self, p):
        assert np.size(p) == self.nparams
        self._params(1, p)

    def get_grads(self):
        """Gradients of the last training step."""
        return self._params(2)

    def forward(self, x):
        """x [B, 2, H, W] float32 CUDA tensor -> x + net(x)."""
        import torch
        B, _, H, W = x.shape
        y = torch.empty_like(x)
        _lib().bart_amd_warmnet_forward(self.h, _fptr(x), B, H, W, _fptr(y))
        _check()
        return y

    def train_step_grad(self, x, dout, lr):
        """Adam step for the output gradient dout of the last forward(x)."""
        B, _, H, W = x.shape
        _lib().bart_amd_warmnet_train_step_grad(self.h, _fptr(x), _fptr(dout), B, H, W, lr)
        _check()

    def train_step(self, x, target, lr):
        B, _, H, W = x.shape
        loss = _lib().bart_amd_warmnet_train_step(self.h, _fptr(x), _fptr(target), B, H, W, lr)
        _check()
        return float(loss)


def to_channels(img, n, scale=1.0):
    """Complex [n*n] device image / scale -> [1, 2, n, n] float32 (real,
    imaginary), packed by the library (bart_amd_warmnet_pack)."""
    import torch
    from . import _dims, _ptr
    out = torch.empty((1, 2, n, n), dtype=torch.float32, device=img.device)
    torch.cuda.synchronize()
    _lib().bart_amd_warmnet_pack(DIMS, _dims([n, n]), _ptr(img.contiguous()), float(scale), _fptr(out))
    _check()
    return out


def from_channels(t, n):
    """[1, 2, n, n] float32 -> complex [n*n] device image."""
    import torch
    return torch.view_as_complex(t[0].permute(1, 2, 0).contiguous()).reshape(n * n).contiguous()


def network_input(op, b, n, lam, d=None):
    """Network input channels [1, 2, n, n] of x1 = alpha d, the best multiple
    of the direction d for the residual of (A^H A + lam I) x = b (d = None:
    d = b, CG's first iterate from zero), scaled by s = max |x1|, and s --
    computed by the library (bart_amd_warmnet_input_dir)."""
    import torch
    from . import _dims, _ptr
    inp = torch.empty((1, 2, n, n), dtype=torch.float32, device=b.device)
    s = ctypes.c_float(0.0)
    torch.cuda.synchronize()
    _lib().bart_amd_warmnet_input_dir(op.h, lam, DIMS, _dims(op.idims), _ptr(b), _ptr(d) if d is not None else None,
                                      _fptr(inp), ctypes.byref(s))
    _check()
    return inp, float(s.value)


def guess(net, op, b, lam, d=None):
    """The learned initial guess x0 = s (x1 / s + net(x1 / s)), x1 as in
    network_input, computed by the library in one call
    (bart_amd_warmnet_guess_dir)."""
    import torch
    from . import _dims, _ptr
    x0 = torch.empty_like(b)
    torch.cuda.synchronize()
    _lib().bart_amd_warmnet_guess_dir(net.h, op.h, lam, DIMS, _dims(op.idims), _ptr(b),
                                      _ptr(d) if d is not None else None, _ptr(x0))
    _check()
    return x0


def guess_dcf(net, op, b, lam, y, w):
    """The learned initial guess along the density-compensated adjoint
    A^H (w y), all of it in the library (bart_amd_warmnet_guess_dcf): the
    k-space weighting, the adjoint, the best multiple, the network."""
    import torch
    from . import _dims, _ptr
    x0 = torch.empty_like(b)
    wd = np.ascontiguousarray(w, np.float32)
    torch.cuda.synchronize()
    _lib().bart_amd_warmnet_guess_dcf(net.h, op.h, lam, DIMS, _dims(op.idims), _ptr(b), _ptr(y),
                                      wd.ctypes.data_as(ctypes.c_void_p), wd.size, _ptr(x0))
    _check()
    return x0


def radial_dcf(R, S):
    """Density compensation of the 2-D radial trajectory of
    inputs.radial_traj (sample m = r + R spoke): the ramp |k| along the
    readout, |r + 1/2 - R/2| (its scale does not matter: x1 takes the best
    multiple)."""
    read = np.abs(np.arange(R, dtype=np.float64) + 0.5 - R / 2.0)
    return np.tile(read, S).astype(np.float32)


def dcf_direction(op, y, w):
    """The density-compensated adjoint A^H (w y) (the coil-combined
    gridding image), the network input direction; y device k-space
    [C][M], w [M]."""
    import torch
    wt = torch.from_numpy(np.ascontiguousarray(w)).to(y.device)
    M = wt.numel()
    return op.adjoint((y.reshape(-1, M) * wt[None]).reshape(-1).contiguous())


# ---------------------------------------------------------------------------
# Training pairs, training, and CG from the network's guess
# ---------------------------------------------------------------------------

def _dims16(*d):
    return list(d) + [1] * (DIMS - len(d))


def make_problem(seed, n=256, C=8, spokes=(32, 96), noise=1e-3, lam=0.01):
    """A seeded 2-D radial CG-SENSE problem: ellipse phantom, Gaussian coil
    maps, golden-angle radial trajectory with a seeded spoke count (readout
    2n), k-space with complex Gaussian noise.  Returns (operator, b = A^H y,
    ground-truth image) with device vectors; the operator carries its
    k-space and density compensation (op.dcf_ksp, op.dcf_w) and the
    density-compensated adjoint (op.dcf_dir, the network input direction of
    the training pairs)."""
    import torch
    from . import inputs as I
    from . import sense_nc_init, traj_array
    rng = np.random.default_rng(seed)
    S = int(rng.integers(spokes[0], spokes[1] + 1))
    R = 2 * n
    traj = I.radial_traj(R, S)
    img = I.ellipse_phantom(n, seed=seed)
    maps = I.gaussian_coils((1, n, n), C, seed=seed)
    op = sense_nc_init(_dims16(n, n, 1, C), _dims16(n, n, 1, C), maps, _dims16(1, R, S, C), _dims16(3, R, S),
                       traj_array(traj))
    x = torch.from_numpy(np.ascontiguousarray(img).reshape(-1)).to("cuda")
    y = op.forward(x)
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    e = (torch.randn(y.numel(), generator=g) + 1j * torch.randn(y.numel(), generator=g)).to(torch.complex64)
    y = (y + (noise * float(y.abs().max()) / np.sqrt(2.0)) * e.to("cuda")).contiguous()
    b = op.adjoint(y)
    op.dcf_ksp, op.dcf_w = y, radial_dcf(R, S)
    op.dcf_dir = dcf_direction(op, y, op.dcf_w)
    return op, b, x


def training_pairs(seeds, n=256, C=8, lam=0.01, maxiter=150, tol=1e-6, chunk=32, log=None):
    """(input, target) pairs: network input from x1, target the CG-SENSE
    solution to relative residual tol (at most maxiter iterations) scaled by
    the same s; problems are solved chunk by chunk with lsqr_conjgrad_batch."""
    import torch
    from . import lsqr_conjgrad_batch
    xs_in, xs_t, its = [], [], []
    for i in range(0, len(seeds), chunk):
        probs = [make_problem(s, n, C, lam=lam) for s in seeds[i:i + chunk]]
        ops = [p[0] for p in probs]
        bs = [p[1] for p in probs]
        sol, st, _ = lsqr_conjgrad_batch(ops, bs, maxiter, lam, tol)
        for (op, b, _), xsol, sti in zip(probs, sol, st):
            inp, s = network_input(op, b, n, lam, op.dcf_dir)
            xs_in.append(inp)
            xs_t.append(to_channels(xsol, n, s))
            its.append(sti.iterations)
        for op in ops:
            op.free()
        torch.cuda.synchronize()
        if log is not None and (i // chunk) % 16 == 15:
            log(f"{i + len(probs)} training pairs")
    return torch.cat(xs_in), torch.cat(xs_t), its


def train(net, inputs, targets, epochs=20, batch=8, lr=1e-3, seed=0, log=None):
    """Adam on mean squared error over shuffled mini-batches; the learning
    rate falls linearly to lr / 10 over the epochs.  Returns the mean loss
    of every epoch."""
    import torch
    rng = np.random.default_rng(seed)
    N = inputs.shape[0]
    hist = []
    for e in range(epochs):
        perm = rng.permutation(N)
        lr_e = lr * (1.0 - 0.9 * e / max(1, epochs - 1))
        tot, cnt = 0.0, 0
        for i in range(0, N - batch + 1, batch):
            idx = perm[i:i + batch]
            ii = torch.from_numpy(idx).to(inputs.device)
            tot += net.train_step(inputs[ii].contiguous(), targets[ii].contiguous(), lr_e)
            cnt += 1
        hist.append(tot / max(1, cnt))
        if log is not None:
            log(f"epoch {e}: loss {hist[-1]:.3e}")
    return hist


def solve_warm_and_cold(net, op, b, n, lam=0.01, tol=1e-4, maxiter=300):
    """CG-SENSE to relative residual tol from zero and from the network's
    guess (zero start as fallback when the guess's residual exceeds ||b||).
    Returns dict with iterations and device times (ms) of both; the warm
    time includes the whole guess (k-space weighting, compensated adjoint,
    best multiple, network)."""
    import torch
    from . import lsqr_conjgrad
    # the operator's iteration graph is captured by its first solve: capture
    # it before either timed solve
    lsqr_conjgrad(op, b, 1, lam, 0.0)
    xc, stc, _ = lsqr_conjgrad(op, b, maxiter, lam, tol)
    def g():
        if hasattr(op, "dcf_ksp"):
            return guess_dcf(net, op, b, lam, op.dcf_ksp, op.dcf_w)
        return guess(net, op, b, lam)
    g()  # warm-up (work buffers)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x0 = g()
    t_guess = (time.perf_counter() - t0) * 1e3
    xw, stw, _ = lsqr_conjgrad(op, b, maxiter, lam, tol, x0=x0, fallback=True)
    diff = float((xw - xc).norm() / xc.norm())
    return dict(cold_it=stc.iterations, warm_it=stw.iterations, cold_ms=stc.time_ms,
                warm_ms=stw.time_ms + t_guess, guess_ms=t_guess, fell_back=bool(stw.fell_back),
                warm_vs_cold=diff, cold_status=stc.status, warm_status=stw.status)


def residual_problems(seeds, n=256, C=8, lam=0.01):
    """Problems kept alive for residual training: (operator, b, network
    input, its scale s, ||b||^2)."""
    out = []
    for sd in seeds:
        op, b, _ = make_problem(sd, n, C, lam=lam)
        inp, s = network_input(op, b, n, lam, op.dcf_dir)
        out.append((op, b, inp, s, float(b.norm()) ** 2))
    return out


def train_residual(net, probs, n, epochs=10, batch=8, lr=2e-4, lam=0.01, seed=0, log=None):
    """Adam on the mean over a batch of ||b - (A^H A + lam I) x0||^2 / ||b||^2,
    x0 = s (network output): the quantity CG's stop test measures.  The
    output gradient -2 s (A^H A + lam I) r / ||b||^2 takes two Toeplitz normal
    operators per image.  Returns the mean loss of every epoch."""
    import torch
    rng = np.random.default_rng(seed)
    hist = []
    for e in range(epochs):
        perm = rng.permutation(len(probs))
        tot, cnt = 0.0, 0
        for i in range(0, len(probs) - batch + 1, batch):
            idx = perm[i:i + batch]
            x = torch.cat([probs[j][2] for j in idx]).contiguous()
            y = net.forward(x)
            douts = []
            for q, j in enumerate(idx):
                op, b, _, s, bn2 = probs[j]
                x0 = (from_channels(y[q:q + 1], n) * s).contiguous()
                r = (b - op.normal(x0) - lam * x0).contiguous()
                g = op.normal(r) + lam * r
                tot += float(r.norm()) ** 2 / bn2
                douts.append(to_channels((-2.0 * s / bn2 / len(idx)) * g, n))
            net.train_step_grad(x, torch.cat(douts).contiguous(), lr)
            cnt += len(idx)
        hist.append(tot / max(1, cnt))
        if log is not None:
            log(f"residual epoch {e}: {hist[-1]:.3e}")
    return hist

