"""GPU: the chained NUFFT for oversampling other than 2 (nufft_create2 with
conf.os != 2, noncart/nufft_chain.c:192-244) against the oracle's
ChainNufft: forward and adjoint to 1e-5 (2-D and 3-D, periodic and
clamped, weights, odd oversampled sizes), the dot test, the reference's
os 1.5 / width 8 test through the command line, and CG-SENSE on the chained
operator against a dense solve of the oracle's normal equations."""
import numpy as np
import pytest

import oracle as O
from bart_amd import inputs as I
from cli_util import dims16, read_raw, run_bart, write_raw

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def d16(*d):
    return list(d) + [1] * (16 - len(d))


def relerr(a, ref):
    a = np.asarray(a, np.complex128).ravel()
    ref = np.asarray(ref, np.complex128).ravel()
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


def crandn(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.complex64).ravel()).to("cuda")


def make(dims, traj, C, w=None, periodic=False, width=6.0, os_=1.5):
    import bart_amd as ba
    conf = ba.NufftConf.defaults()
    conf.os = os_
    conf.width = width
    conf.periodic = periodic
    M = traj.shape[0]
    return ba.nufft_create2(d16(1, M, 1, C), d16(*dims, C), d16(3, M, 1), np.ascontiguousarray(traj),
                            wgh_dims=d16(1, M, 1) if w is not None else None, weights=w, conf=conf)


CASES = [
    ((32, 32, 1), "radial", 1.5, 8.0, True, False),
    ((32, 32, 1), "radial", 1.25, 6.0, False, True),
    ((30, 30, 1), "radial", 1.5, 6.0, True, True),     # oversampled size 45: cells of one point
    ((16, 16, 16), "koosh", 1.5, 8.0, False, False),
    ((16, 16, 16), "koosh", 1.25, 5.0, True, True),
]


@pytest.mark.parametrize("dims,kind,os_,width,periodic,weights", CASES)
def test_chain_vs_oracle(dims, kind, os_, width, periodic, weights):
    C = 3
    traj = I.phyllotaxis_kooshball(spokes=80, readout=32) if kind == "koosh" else I.radial_traj(64, 21)
    M = traj.shape[0]
    w = (0.5 + np.random.default_rng(2).random(M)).astype(np.complex64) if weights else None
    ref = O.ChainNufft(dims, traj, C, weights=w, periodic=periodic, width=width, os_=os_)
    op = make(dims, traj, C, w, periodic, width, os_)
    x = crandn((C,) + dims[::-1], 3)
    y = crandn((C, M), 4)
    f = op.forward(dev(x)).cpu().numpy()
    a = op.adjoint(dev(y)).cpu().numpy()
    e_f, e_a = relerr(f, ref.forward(x)), relerr(a, ref.adjoint(y))
    print(f"{dims} os {os_} w {width}: forward {e_f:.1e} adjoint {e_a:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5
    lhs = np.vdot(f.astype(np.complex128), y.ravel())
    rhs = np.vdot(x.ravel().astype(np.complex128), a)
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
    # repeated applies are bitwise equal (the adjoint gathers in fixed order)
    assert np.array_equal(op.adjoint(dev(y)).cpu().numpy(), a)
    op.free()


def test_chain_full_size_radial_vs_oracle():
    # 2-D golden-angle radial 256^2, 201 x 512, os 1.5, 12 coils: the dense
    # centre tiles of the adjoint are split over several work-groups (slices
    # of 1024 samples summed in order) and the coils over two chunks of 8
    n, C, R, S = 256, 12, 512, 201
    traj = I.radial_traj(R, S)
    ref = O.ChainNufft((n, n, 1), traj, C, os_=1.5, width=6.0)
    op = make((n, n, 1), traj, C, os_=1.5, width=6.0)
    x = crandn((C, 1, n, n), 5)
    y = crandn((C, R * S), 6)
    a = op.adjoint(dev(y)).cpu().numpy()
    e_f, e_a = relerr(op.forward(dev(x)).cpu().numpy(), ref.forward(x)), relerr(a, ref.adjoint(y))
    print(f"256^2 radial os 1.5, 12 coils: forward {e_f:.1e} adjoint {e_a:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5
    assert np.array_equal(op.adjoint(dev(y)).cpu().numpy(), a)
    op.free()


def test_cli_nufft_os_equals_os2(tmp_path):
    # tests/nufft.mk:305-323: periodic forward and adjoint with os 1.5 and
    # width 8 agree with os 2 and width 6 to 1e-5
    run_bart(["traj", "-r", "-x64", "-y64", "t"], tmp_path)
    n = 64
    write_raw(str(tmp_path / "x"), dims16(n, n, 1), I.ellipse_phantom(n, seed=9))
    run_bart(["nufft", "-P", "-r", "t", "x", "k1"], tmp_path)
    run_bart(["nufft", "-P", "-w8", "-o1.5", "-t", "t", "x", "k2"], tmp_path)
    _, k1 = read_raw(str(tmp_path / "k1"))
    _, k2 = read_raw(str(tmp_path / "k2"))
    e_f = relerr(k2, k1)
    run_bart(["nufft", "-a", "-P", "-r", "t", "k1", "r1"], tmp_path)
    run_bart(["nufft", "-a", "-P", "-w8", "-o1.5", "-t", "t", "k1", "r2"], tmp_path)
    _, r1 = read_raw(str(tmp_path / "r1"))
    _, r2 = read_raw(str(tmp_path / "r2"))
    e_a = relerr(r2, r1)
    print(f"os 1.5 / w 8 vs os 2: forward {e_f:.1e}, adjoint {e_a:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5


def test_chain_sense_cgsense():
    import bart_amd as ba
    n, C = 16, 3
    traj = I.radial_traj(32, 15)
    M = traj.shape[0]
    maps = I.gaussian_coils((1, n, n), C, seed=5)
    conf = ba.NufftConf.defaults()
    conf.os = 1.5
    conf.width = 8.0
    op = ba.sense_nc_init(d16(n, n, 1, C), d16(n, n, 1, C), maps, d16(1, M, 1, C), d16(3, M, 1),
                          np.ascontiguousarray(traj), conf=conf)
    ref = O.ChainNufft((n, n, 1), traj, C, width=8.0, os_=1.5)
    m = maps.reshape(C, 1, n, n)

    def normal(x):
        return np.sum(ref.adjoint(ref.forward(x.reshape(1, 1, n, n) * m)) * np.conj(m), axis=0)

    x = crandn((1, n, n), 6)
    e_n = relerr(op.normal(dev(x)).cpu().numpy(), normal(x))
    assert e_n < 1e-5
    lam = 0.05
    A = np.zeros((n * n, n * n), np.complex128)
    for j in range(n * n):
        e = np.zeros(n * n, np.complex64)
        e[j] = 1
        A[:, j] = normal(e).ravel()
    y = crandn((C, M), 7)
    b = op.adjoint(dev(y))
    xs = np.linalg.solve(A + lam * np.eye(n * n), b.cpu().numpy().ravel().astype(np.complex128))
    xg, st, _ = ba.lsqr_conjgrad(op, b, 300, lam, 1e-6)
    e_cg = relerr(xg.cpu().numpy(), xs)
    print(f"chained SENSE normal {e_n:.1e}; CG-SENSE {st.iterations} iterations, vs dense solve {e_cg:.1e}")
    assert st.status == 2 and e_cg < 1e-4
    op.free()


def test_non_decomposed_grid_vs_oracle():
    # decomp = false (nufft -1): the image is the oversampled grid itself
    import bart_amd as ba
    n, C = 64, 2
    traj = I.radial_traj(64, 31)
    M = traj.shape[0]
    w = (0.5 + np.random.default_rng(3).random(M)).astype(np.complex64)
    conf = ba.NufftConf.defaults(decomp=False, toeplitz=False)
    op = ba.nufft_create2(d16(1, M, 1, C), d16(n, n, 1, C), d16(3, M, 1), np.ascontiguousarray(traj),
                          wgh_dims=d16(1, M, 1), weights=w, conf=conf)
    ref = O.ChainNufft((n, n, 1), traj, C, weights=w, width=6.0, os_=2.0, pad=False)
    x = crandn((C, 1, n, n), 12)
    y = crandn((C, M), 13)
    e_f = relerr(op.forward(dev(x)).cpu().numpy(), ref.forward(x))
    e_a = relerr(op.adjoint(dev(y)).cpu().numpy(), ref.adjoint(y))
    print(f"non-decomposed: forward {e_f:.1e} adjoint {e_a:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5
    op.free()
    with pytest.raises(ba.BartError, match="Toeplitz"):
        ba.nufft_create2(d16(1, M, 1, C), d16(n, n, 1, C), d16(3, M, 1), np.ascontiguousarray(traj),
                         conf=ba.NufftConf.defaults(decomp=False))


def test_cli_nufft_over(tmp_path):
    # tests/nufft.mk:336-347 (test-nufft-over): nufft -1 of the zero-padded
    # image equals nufft of the image; nufft -a -1 cropped equals nufft -a
    run_bart(["traj", "-r", "-x128", "-y128", "t"], tmp_path)
    n = 128
    x = I.ellipse_phantom(n, seed=14).reshape(n, n)
    x2 = np.zeros((2 * n, 2 * n), np.complex64)
    x2[n // 2:n // 2 + n, n // 2:n // 2 + n] = x
    write_raw(str(tmp_path / "x"), dims16(n, n, 1), x)
    write_raw(str(tmp_path / "x2"), dims16(2 * n, 2 * n, 1), x2)
    run_bart(["nufft", "-r", "t", "x", "k1"], tmp_path)
    run_bart(["nufft", "-1", "-r", "t", "x2", "k2"], tmp_path)
    _, k1 = read_raw(str(tmp_path / "k1"))
    _, k2 = read_raw(str(tmp_path / "k2"))
    e_f = relerr(k2, k1)
    run_bart(["nufft", "-a", "t", "k1", "r1"], tmp_path)
    run_bart(["nufft", "-a", "-1", "-r", "t", "k1", "r2a"], tmp_path)
    d1, r1 = read_raw(str(tmp_path / "r1"))
    d2, r2a = read_raw(str(tmp_path / "r2a"))
    assert d1 == dims16(n, n, 1) and d2 == dims16(2 * n, 2 * n, 1)
    r2 = r2a.reshape(2 * n, 2 * n)[n // 2:n // 2 + n, n // 2:n // 2 + n]
    e_a = relerr(r2, r1)
    print(f"nufft -1 of the padded image vs nufft: forward {e_f:.1e}, adjoint cropped {e_a:.1e}")
    assert e_f < 1e-6 and e_a < 1e-6
