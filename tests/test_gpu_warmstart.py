"""GPU: the warm-start network (bart_amd_warmnet_*) against the FP64 oracle
network of oracle/cnn.py evaluated with the same bfloat16 rounding points:
forward output, loss and every parameter gradient of one training step
(backpropagation through the matrix-core kernels), and that training lowers
the loss."""
import numpy as np
import pytest

from oracle import cnn as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")


def relerr(a, ref):
    a = np.asarray(a, np.float64)
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


def test_warmnet_forward_loss_and_gradients():
    from bart_amd.warmstart import WarmNet
    ch = 8
    net = WarmNet(ch, seed=3)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 2, 8, 16)).astype(np.float32)
    t = rng.standard_normal((2, 2, 8, 16)).astype(np.float32)
    y = net.forward(dev(x)).cpu().numpy()
    ref, _ = R.net_forward(p, x, ch)
    assert relerr(y, ref) < 1e-4
    g_ref, loss_ref = R.net_grads(p, x, t, ch)
    loss = net.train_step(dev(x), dev(t), 1e-3)
    assert abs(loss - loss_ref) < 1e-4 * loss_ref
    g = net.get_grads()
    print(f"forward rel err {relerr(y, ref):.1e}, gradient rel err {relerr(g, g_ref):.1e}")
    assert relerr(g, g_ref) < 1e-3
    # unrounded FP64 gradients: the bfloat16 operands show at the 1e-2 level
    g64, _ = R.net_grads(p, x, t, ch, rounding=False)
    assert relerr(g, g64) < 5e-2
    # the first Adam step moves every parameter with a non-zero gradient by lr
    d = net.get_params() - p
    nz = np.abs(g) > 1e-6
    assert np.allclose(np.abs(d[nz]), 1e-3, rtol=2e-2)
    net.free()


def test_warmnet_params_roundtrip():
    from bart_amd.warmstart import WarmNet
    a = WarmNet(16, seed=1)
    b = WarmNet(16, seed=2)
    pa = a.get_params()
    assert not np.array_equal(pa, b.get_params())
    b.set_params(pa)
    x = dev(np.random.default_rng(0).standard_normal((1, 2, 12, 64)))
    assert np.array_equal(a.forward(x).cpu().numpy(), b.forward(x).cpu().numpy())


def test_warmnet_training_lowers_loss():
    # denoising of smooth random fields: the loss falls well below its start
    from bart_amd.warmstart import WarmNet
    rng = np.random.default_rng(9)
    n = 32
    yy, xx = np.mgrid[0:n, 0:n] / n
    clean = np.stack([np.stack([np.sin(2 * np.pi * (rng.random() * 2 * xx + rng.random() * 2 * yy + rng.random()))
                                for _ in range(2)]) for _ in range(8)]).astype(np.float32)
    noisy = clean + 0.3 * rng.standard_normal(clean.shape).astype(np.float32)
    net = WarmNet(16, seed=0)
    xd, td = dev(noisy), dev(clean)
    losses = [net.train_step(xd, td, 2e-3) for _ in range(150)]
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.5 * losses[0]


def test_warm_start_pipeline_and_fallback():
    # training pairs from the solver, a few training steps, CG to 1e-4 from
    # the guess and from zero; a network scaled far off falls back to zero
    from bart_amd import warmstart as WS
    n = 64
    inp, tgt, its = WS.training_pairs(list(range(8)), n=n, C=4, maxiter=60, chunk=4)
    assert inp.shape == (8, 2, n, n) and tgt.shape == (8, 2, n, n)
    net = WS.WarmNet(8, seed=0)
    hist = WS.train(net, inp, tgt, epochs=3, batch=4, lr=1e-3)
    assert np.isfinite(hist).all()
    op, b, _ = WS.make_problem(1000, n=n, C=4)
    r = WS.solve_warm_and_cold(net, op, b, n)
    assert r["cold_status"] == 2 and r["warm_status"] == 2
    # both meet ||r|| <= 1e-4 ||b||; the solutions agree to the problem's conditioning
    assert r["warm_vs_cold"] < 0.1
    p = net.get_params()
    net.set_params(p * 50.0)
    r2 = WS.solve_warm_and_cold(net, op, b, n)
    assert r2["fell_back"] and r2["warm_it"] == r2["cold_it"]
    op.free()


def test_train_step_with_given_output_gradient():
    # the MSE step and the step fed the MSE's output gradient by hand agree
    from bart_amd.warmstart import WarmNet
    rng = np.random.default_rng(12)
    x = dev(rng.standard_normal((2, 2, 8, 64)))
    t = dev(rng.standard_normal((2, 2, 8, 64)))
    a = WarmNet(8, seed=5)
    b = WarmNet(8, seed=5)
    a.train_step(x, t, 1e-3)
    y = b.forward(x)
    dout = (2.0 * (y - t) / y.numel()).contiguous()
    b.train_step_grad(x, dout, 1e-3)
    assert np.allclose(a.get_params(), b.get_params(), rtol=0, atol=1e-7)
    assert np.allclose(a.get_grads(), b.get_grads(), rtol=1e-6, atol=1e-12)


def test_network_input_and_guess_entries():
    # bart_amd_warmnet_input / _guess: x1 = alpha b is CG's first iterate from
    # zero (bitwise the 1-iteration solve), s = max |x1|, the channels of
    # x1 / s, and x0 = s (x1 / s + net(x1 / s)); host pointers as well
    import bart_amd as ba
    from bart_amd import warmstart as WS
    n = 64
    op, b, _ = WS.make_problem(7, n=n, C=4)
    lam = 0.01
    x1, _, _ = ba.lsqr_conjgrad(op, b, 1, lam, 0.0)
    inp, s = WS.network_input(op, b, n, lam)
    s_ref = float(x1.abs().max())
    assert abs(s / s_ref - 1) < 1e-6
    ch = torch.stack([x1.real, x1.imag]).reshape(1, 2, n, n) / s
    assert float((inp - ch).abs().max()) <= 1e-6 * float(ch.abs().max())
    net = WS.WarmNet(8, seed=2)
    x0 = WS.guess(net, op, b, lam)
    y = net.forward(inp)
    x0_ref = torch.complex(y[0, 0], y[0, 1]).reshape(-1) * s
    assert float((x0 - x0_ref).abs().max()) <= 1e-6 * float(x0_ref.abs().max())
    # host vectors through the same entry
    bh = b.cpu().numpy()
    xh = np.zeros_like(bh)
    ba.lib().bart_amd_warmnet_guess(net.h, op.h, lam, ba.DIMS, ba._dims(op.idims), ba._ptr(bh), ba._ptr(xh))
    ba._check()
    assert np.array_equal(xh, x0.cpu().numpy())
    # the training target packing
    t = WS.to_channels(x1, n, s)
    assert torch.equal(t, inp) or float((t - inp).abs().max()) <= 1e-6
    net.free()
    op.free()


def test_network_input_along_a_direction():
    # bart_amd_warmnet_input_dir / _guess_dir with d the density-compensated
    # adjoint: x1 = alpha d, alpha = d^H b / d^H (A^H A d + lambda d) (the
    # best multiple of d for the residual, in complex128 here), s = max |x1|
    import bart_amd as ba
    from bart_amd import warmstart as WS
    n, lam = 64, 0.01
    op, b, _ = WS.make_problem(11, n=n, C=4)
    d = op.dcf_dir
    inp, s = WS.network_input(op, b, n, lam, d)
    d64 = d.to(torch.complex128)
    Nd = (op.normal(d) + lam * d).to(torch.complex128)
    alpha = torch.vdot(d64, b.to(torch.complex128)) / torch.vdot(d64, Nd)
    x1 = alpha * d64
    s_ref = float(x1.abs().max())
    assert abs(s / s_ref - 1) < 1e-5
    ch = torch.stack([x1.real, x1.imag]).reshape(1, 2, n, n).float() / s
    assert float((inp - ch).abs().max()) <= 1e-5 * float(ch.abs().max())
    # the best multiple: the residual grows when alpha is perturbed
    def res(a):
        x = (a * d64).to(torch.complex64)
        return float((b - op.normal(x) - lam * x).norm())
    assert res(alpha) < min(res(alpha * 1.05), res(alpha * 0.95), res(alpha * 1j ** 0.1))
    # d = None is the first CG iterate (the plain entry)
    inp0, s0 = WS.network_input(op, b, n, lam)
    inp1, s1 = WS.network_input(op, b, n, lam, b)
    assert s0 == s1 and torch.equal(inp0, inp1)
    net = WS.WarmNet(8, seed=3)
    x0 = WS.guess(net, op, b, lam, d)
    y = net.forward(inp)
    x0_ref = torch.complex(y[0, 0], y[0, 1]).reshape(-1) * s
    assert float((x0 - x0_ref).abs().max()) <= 1e-6 * float(x0_ref.abs().max())
    # the whole guess in the library (k-space weighting and adjoint
    # included), device and host k-space
    x0d = WS.guess_dcf(net, op, b, lam, op.dcf_ksp, op.dcf_w)
    assert float((x0d - x0).abs().max()) <= 1e-6 * float(x0.abs().max())
    x0h = WS.guess_dcf(net, op, b, lam, op.dcf_ksp.cpu().numpy(), op.dcf_w)
    assert torch.equal(x0h, x0d)
    net.free()
    op.free()


def test_warm_start_full_size():
    # the learned warm start at its workload: 2-D radial 256^2, 8 coils,
    # network of 48 channels.  Training pairs from the batched solver (CG to
    # relative residual 1e-6), the network's forward output, loss and every
    # parameter gradient against the oracle network on one full image, a
    # few epochs of training, and CG to 1e-4 from the guess and from zero
    from bart_amd import warmstart as WS
    n, ch = 256, 48
    inp, tgt, its = WS.training_pairs(list(range(100, 108)), n=n, C=8, maxiter=150, tol=1e-6, chunk=8)
    assert inp.shape == (8, 2, n, n) and tgt.shape == (8, 2, n, n)
    print("target CG iterations", its)
    net = WS.WarmNet(ch, seed=0)
    p = net.get_params().astype(np.float64)
    x, t = inp[:1].contiguous(), tgt[:1].contiguous()
    xh, th = x.cpu().numpy(), t.cpu().numpy()
    y = net.forward(x).cpu().numpy()
    ref, _ = R.net_forward(p, xh, ch)
    e_fwd = relerr(y, ref)
    g_ref, loss_ref = R.net_grads(p, xh, th, ch)
    loss = net.train_step(x, t, 1e-3)
    e_grad = relerr(net.get_grads(), g_ref)
    print(f"256^2 x 48 channels: forward rel err {e_fwd:.1e}, loss {loss:.4e} (oracle {loss_ref:.4e}), "
          f"gradient rel err {e_grad:.1e}")
    assert e_fwd < 1e-4
    assert abs(loss - loss_ref) < 1e-4 * loss_ref
    assert e_grad < 1e-3
    hist = WS.train(net, inp, tgt, epochs=6, batch=4, lr=1e-3)
    print("epoch losses", " ".join(f"{h:.3e}" for h in hist))
    assert np.isfinite(hist).all() and hist[-1] < hist[0]
    op, b, _ = WS.make_problem(2000, n=n, C=8)
    r = WS.solve_warm_and_cold(net, op, b, n)
    print(r)
    assert r["cold_status"] == 2 and r["warm_status"] == 2
    assert r["warm_vs_cold"] < 0.1
    op.free()
    net.free()
