"""GPU: the rest of BART's NUFFT boundary through the C ABI --
nufft_create_normal, nufft_update_traj, nufft_get_psf2 / nufft_update_psf2
(byte strides, host and device memory), compute_psf2 -- and the per-handle
stream of bart_amd_linop_set_stream."""
import numpy as np
import pytest

import bart_amd as ba
import oracle as O
from bart_amd import inputs as I
from conftest import crandn, relerr

pytestmark = pytest.mark.gpu


def dims16(*d):
    return list(d) + [1] * (16 - len(d))


def make(img, C, traj, R, S, weights=None, **conf):
    X, Y, Z = img
    c = ba.NufftConf.defaults(**conf)
    return ba.nufft_create2(dims16(1, R, S, C), dims16(X, Y, Z, C), dims16(3, R, S), ba.traj_array(traj),
                            dims16(1, R, S) if weights is not None else None, weights, c)


@pytest.mark.parametrize("img,R,S", [((64, 64, 1), 128, 21), ((32, 32, 32), 64, 90)])
def test_create_normal_from_psf(img, R, S):
    # nufft.c:1285-1312: the normal operator of a PSF alone equals the
    # Toeplitz normal of the NUFFT the PSF came from; forward is an error
    X, Y, Z = img
    C = 2
    traj = I.radial_traj(R, S) if Z == 1 else I.phyllotaxis_kooshball(spokes=S, readout=R)
    op = make(img, C, traj, R, S, toeplitz=True)
    psf = op.get_psf()
    nop = ba.nufft_create_normal(dims16(X, Y, Z, C), op.psf_dims(), psf)
    x = crandn((C, Z, Y, X), 301)
    assert np.array_equal(nop.normal(x), op.normal(x))
    with pytest.raises(ba.BartError, match="normal operator only"):
        nop.forward(x)
    # created without a PSF, then given one
    nop2 = ba.nufft_create_normal(dims16(X, Y, Z, C), op.psf_dims())
    nop2.update_psf(psf)
    assert np.array_equal(nop2.normal(x), op.normal(x))
    for o in (op, nop, nop2):
        o.free()


def test_update_traj_matches_fresh_operator():
    # nufft.c:2041-2054: a new trajectory and weights of the same size give
    # the operator built on them from the start (plan and PSF rebuilt)
    n, C, R, S = 64, 3, 128, 25
    ta = I.radial_traj(R, S)
    tb = I.radial_traj(R, S, golden=False)
    wb = (0.5 + np.random.default_rng(3).random(R * S)).astype(np.complex64)
    op = make((n, n, 1), C, ta, R, S, toeplitz=True)
    x = crandn((C, 1, n, n), 302)
    y = crandn((C, R * S), 303)
    op.forward(x)
    op.normal(x)
    op.update_traj(dims16(3, R, S), ba.traj_array(tb), weights=wb)
    fresh = make((n, n, 1), C, tb, R, S, weights=wb, toeplitz=True)
    assert np.array_equal(op.forward(x), fresh.forward(x))
    assert np.array_equal(op.adjoint(y), fresh.adjoint(y))
    assert np.array_equal(op.get_psf(), fresh.get_psf())
    assert np.array_equal(op.normal(x), fresh.normal(x))
    ref = O.Nufft((n, n, 1), tb, C, weights=wb, toeplitz=True)
    assert relerr(op.normal(x).reshape(C, 1, n, n), ref.normal(x)) < 1e-5
    with pytest.raises(ba.BartError):
        op.update_traj(dims16(3, R, S - 1), ba.traj_array(tb[:R * (S - 1)]))
    op.free()
    fresh.free()


def test_update_traj_of_sense_operator_and_cg():
    # a SENSE operator after a trajectory update solves as a fresh one
    # (the captured CG iteration is rebuilt)
    n, C, R, S = 64, 4, 128, 31
    maps = I.gaussian_coils((1, n, n), C, seed=5)
    ta = I.radial_traj(R, S)
    tb = I.radial_traj(R, S, golden=False)
    mk = lambda t: ba.sense_nc_init(dims16(n, n, 1, C), dims16(n, n, 1, C), maps, dims16(1, R, S, C),
                                    dims16(3, R, S), ba.traj_array(t))
    op, fresh = mk(ta), mk(tb)
    b = crandn((1, n, n), 304)
    ba.lsqr_conjgrad(op, b, 5, 0.01, 0.0)
    op.update_traj(dims16(3, R, S), ba.traj_array(tb))
    x1, _, _ = ba.lsqr_conjgrad(op, b, 10, 0.01, 0.0)
    x2, _, _ = ba.lsqr_conjgrad(fresh, b, 10, 0.01, 0.0)
    assert np.array_equal(x1, x2)
    op.free()
    fresh.free()


def test_psf2_strides_host_and_device():
    # nufft_get_psf2 / nufft_update_psf2 with byte strides: the PSF written
    # into every other element of a larger buffer and read back from it
    torch = pytest.importorskip("torch")
    n, R, S = 32, 64, 11
    op = make((n, n, 1), 1, I.radial_traj(R, S), R, S, toeplitz=True)
    d = op.psf_dims()
    dense = op.get_psf()
    cnt = int(np.prod(d))
    strides = []
    acc = 16
    for v in d:
        strides.append(acc)
        acc *= v
    big = np.zeros(2 * cnt, np.complex64)
    op.get_psf2(big, strides)
    assert np.array_equal(big[::2], dense) and not big[1::2].any()
    bigd = torch.zeros(2 * cnt, dtype=torch.complex64, device="cuda")
    op.get_psf2(bigd, strides)
    assert np.array_equal(bigd.cpu().numpy()[::2], dense)
    x = crandn((1, 1, n, n), 305)
    y0 = op.normal(x)
    op.update_psf2(3.0 * big, strides)
    assert relerr(op.normal(x), 3.0 * y0) < 1e-6
    op.update_psf2((2.0 * bigd).contiguous(), strides)
    assert relerr(op.normal(x), 2.0 * y0) < 1e-6
    op.free()


def test_compute_psf2_matches_operator_psf():
    # nufft.c:727-776: the PSF of compute_psf2 is the one the Toeplitz
    # operator builds (periodic 2x-grid adjoint), and the oracle's
    n, R, S = 48, 96, 13
    traj = I.radial_traj(R, S)
    w = (0.5 + np.random.default_rng(4).random(R * S)).astype(np.complex64)
    op = make((n, n, 1), 1, traj, R, S, weights=w, toeplitz=True)
    p = ba.compute_psf2(op.psf_dims(), dims16(3, R, S), ba.traj_array(traj), weights=w)
    assert np.array_equal(p, op.get_psf())
    ref = O.compute_psf((n, n, 1), traj, w)
    assert relerr(p.reshape(ref.shape), ref) < 1e-5
    op.free()


def test_stream_setter_orders_with_torch_stream():
    # applies on a torch side stream: inputs written there, outputs read
    # there, no host synchronisation in between
    torch = pytest.importorskip("torch")
    n, C, R, S = 64, 2, 128, 21
    traj = I.radial_traj(R, S)
    op = make((n, n, 1), C, traj, R, S, toeplitz=True)
    x = crandn((C, 1, n, n), 306)
    ref_k = op.forward(x)
    ref_n = op.normal(x)
    side = torch.cuda.Stream()
    xd = torch.from_numpy(x.reshape(-1)).to("cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        x2 = (xd * 2.0).contiguous()
        k = op.forward(x2)
        k2 = (k * 0.5).contiguous()
        nn = op.normal(x2)
        nn2 = (nn * 0.5).contiguous()
    side.synchronize()
    assert np.array_equal(k2.cpu().numpy(), ref_k)
    assert relerr(nn2.cpu().numpy(), ref_n) < 1e-6
    op.free()
