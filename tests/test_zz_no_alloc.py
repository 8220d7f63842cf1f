"""lsqr_conjgrad_noalloc (the allocation-free CG solve, working memory in a
static array of the library) against lsqr_conjgrad through the C ABI: the
same solutions, iteration counts and residual traces bit for bit, on two
calls in a row (the second smaller, so stale working memory of the first
would show), from zero and from a warm start with fallback; and the limits
are errors, not faults."""
import numpy as np
import pytest

import bart_amd as ba
from bart_amd import inputs as I

pytestmark = pytest.mark.gpu


def d16(*d):
    return list(d) + [1] * (16 - len(d))


def sense(n, C, R, S, seed, toeplitz=True):
    traj = I.radial_traj(R, S)
    maps = I.gaussian_coils((1, n, n), C, seed=seed)
    conf = ba.NufftConf.defaults(toeplitz=toeplitz)
    return ba.sense_nc_init(d16(n, n, 1, C), d16(n, n, 1, C), maps, d16(1, R, S, C), d16(3, R, S),
                            ba.traj_array(traj), conf=conf)


def rhs(op, n, seed):
    import torch
    img = I.ellipse_phantom(n, seed=seed)
    y = op.forward(torch.from_numpy(img.reshape(-1).copy()).cuda())
    return op.adjoint(y)


@pytest.mark.parametrize("tol", [0.0, 1e-4])
def test_noalloc_equals_lsqr_conjgrad_twice(tol):
    import torch
    for n, C, R, S, seed, it in [(96, 6, 192, 61, 21, 30), (48, 3, 96, 29, 22, 20)]:
        op = sense(n, C, R, S, seed)
        b = rhs(op, n, seed)
        torch.cuda.synchronize()
        xr, sr, tr = ba.lsqr_conjgrad(op, b, it, 0.01, tol)
        xn, sn, tn = ba.lsqr_conjgrad_noalloc(op, b, it, 0.01, tol)
        torch.cuda.synchronize()
        assert sn.iterations == sr.iterations and sn.status == sr.status
        assert sn.resid == sr.resid and sn.bnorm == sr.bnorm
        assert np.array_equal(tn, tr)
        assert torch.equal(xn, xr), f"{n}^2: solutions differ"
        op.free()


def test_noalloc_warm_start_fallback_and_explicit_normal():
    import torch
    n = 64
    op = sense(n, 4, 128, 41, 23)
    b = rhs(op, n, 23)
    g = torch.Generator(device="cpu").manual_seed(5)
    x0 = (torch.randn(n * n, generator=g) + 1j * torch.randn(n * n, generator=g)).to(torch.complex64).cuda()
    for x0v, fb in [(0.9 * b, True), (1e3 * x0, True), (x0, False)]:
        xr, sr, tr = ba.lsqr_conjgrad(op, b, 12, 0.01, 0.0, x0=x0v, fallback=fb)
        xn, sn, tn = ba.lsqr_conjgrad_noalloc(op, b, 12, 0.01, 0.0, x0=x0v, fallback=fb)
        torch.cuda.synchronize()
        assert sn.fell_back == sr.fell_back and sn.iterations == sr.iterations
        assert np.array_equal(tn, tr) and torch.equal(xn, xr)
    op.free()
    # adjoint(forward(x)) as the normal operator: its k-space buffer comes
    # from an earlier lsqr_conjgrad call
    ope = sense(n, 4, 128, 41, 23, toeplitz=False)
    xr, sr, tr = ba.lsqr_conjgrad(ope, b, 8, 0.01, 0.0)
    xn, sn, tn = ba.lsqr_conjgrad_noalloc(ope, b, 8, 0.01, 0.0)
    torch.cuda.synchronize()
    assert np.array_equal(tn, tr) and torch.equal(xn, xr)
    ope.free()


def test_noalloc_limits_are_errors():
    n = 32
    op = sense(n, 2, 64, 21, 24)
    b = rhs(op, n, 24)
    with pytest.raises(ba.BartError, match="iterations exceed the maximum"):
        ba.lsqr_conjgrad_noalloc(op, b, 5000, 0.01, 0.0)
    ope = sense(n, 2, 64, 21, 24, toeplitz=False)
    with pytest.raises(ba.BartError, match="k-space buffer"):
        ba.lsqr_conjgrad_noalloc(ope, b, 5, 0.01, 0.0)
    ope.free()
    # a host vector: rejected by the C entry point itself
    xh = np.zeros(n * n, np.complex64)
    it = ba.lib().lsqr_conjgrad_noalloc(op.h, 5, 0.01, 0.0, ba.DIMS, ba._dims(op.idims), ba._ptr(xh), ba._ptr(b),
                                        False, False, None, None)
    assert it == -1
    with pytest.raises(ba.BartError, match="device pointers"):
        ba._check()
    op.free()
