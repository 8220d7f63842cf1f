"""GPU: the NUFFT with a subspace basis (nufft_create2 / sense_nc_init with
basis dims [1 1 1 1 1 T K], noncart/nufft.c:470-505, 1018-1036, 1450-1487,
1546-1625) against the oracle's BasisNufft: forward, adjoint and the
Toeplitz normal operator with the K x K PSF to 1e-5, 2-D and 3-D, per-frame
and shared trajectories, weights; the dot test; the SENSE operator and
CG-SENSE against a dense solve of the oracle's normal equations."""
import numpy as np
import pytest

import oracle as O
from bart_amd import inputs as I

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DIMS = 16


def d16(*d):
    return list(d) + [1] * (DIMS - len(d))


def relerr(a, ref):
    a = np.asarray(a, np.complex128).ravel()
    ref = np.asarray(ref, np.complex128).ravel()
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


def crandn(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def case(dims, T, K, spf, R, seed, per_frame=True, weights=False, kooshball=False):
    rng = np.random.default_rng(seed)
    if kooshball:
        traj = I.phyllotaxis_kooshball(spokes=T * spf, readout=R).reshape(T, R * spf, 3)
    else:
        traj = I.radial_traj(R, T * spf).reshape(T, R * spf, 3)
    if not per_frame:
        traj = traj[:1]
    B = (rng.standard_normal((T, K)) + 1j * rng.standard_normal((T, K))).astype(np.complex64)
    w = (0.5 + rng.random((traj.shape[0], R * spf))).astype(np.complex64) if weights else None
    return traj, B, w


def make_op(dims, traj, B, C, w=None, toeplitz=True, real=False):
    import bart_amd as ba
    T, K = B.shape
    Tt, Mf = traj.shape[0], traj.shape[1]
    conf = ba.NufftConf.defaults()
    conf.toeplitz = toeplitz
    conf.real = real
    op = ba.nufft_create2(d16(1, Mf, 1, C, 1, T), d16(*dims, C, 1, 1, K), d16(3, Mf, 1, 1, 1, Tt),
                          np.ascontiguousarray(traj), wgh_dims=d16(1, Mf, 1, 1, 1, Tt) if w is not None else None,
                          weights=w, conf=conf, bas_dims=d16(1, 1, 1, 1, 1, T, K),
                          basis=np.ascontiguousarray(B.T))
    return op


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.complex64).ravel()).to("cuda")


@pytest.mark.parametrize("per_frame,weights", [(True, False), (True, True), (False, True)])
def test_basis_nufft_2d_vs_oracle(per_frame, weights):
    n, C = 32, 2
    traj, B, w = case((n, n, 1), T=3, K=2, spf=8, R=64, seed=1, per_frame=per_frame, weights=weights)
    T, K = B.shape
    ref = O.BasisNufft((n, n, 1), traj, B, C, weights=w)
    op = make_op((n, n, 1), traj, B, C, w)
    x = crandn((K, C, 1, n, n), 2)
    y = op.forward(dev(x)).cpu().numpy()
    e_f = relerr(y, ref.forward(x))
    yk = crandn((T, C, traj.shape[1]), 3)
    e_a = relerr(op.adjoint(dev(yk)).cpu().numpy(), ref.adjoint(yk))
    e_n = relerr(op.normal(dev(x)).cpu().numpy(), ref.normal(x))
    print(f"forward {e_f:.1e} adjoint {e_a:.1e} normal {e_n:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5 and e_n < 1e-5
    # dot test of the device operator
    ax = op.forward(dev(x)).cpu().numpy().astype(np.complex128)
    ahy = op.adjoint(dev(yk)).cpu().numpy().astype(np.complex128)
    lhs, rhs = np.vdot(ax, yk.ravel()), np.vdot(x.ravel().astype(np.complex128), ahy)
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
    op.free()


def test_basis_nufft_3d_slice_path_vs_oracle():
    n, C = 16, 2
    traj, B, w = case((n, n, n), T=2, K=2, spf=60, R=32, seed=5, weights=True, kooshball=True)
    T, K = B.shape
    ref = O.BasisNufft((n, n, n), traj, B, C, weights=w)
    op = make_op((n, n, n), traj, B, C, w)
    x = crandn((K, C, n, n, n), 6)
    e_f = relerr(op.forward(dev(x)).cpu().numpy(), ref.forward(x))
    yk = crandn((T, C, traj.shape[1]), 7)
    e_a = relerr(op.adjoint(dev(yk)).cpu().numpy(), ref.adjoint(yk))
    e_n = relerr(op.normal(dev(x)).cpu().numpy(), ref.normal(x))
    print(f"3-D forward {e_f:.1e} adjoint {e_a:.1e} normal {e_n:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5 and e_n < 1e-5
    op.free()


def test_basis_explicit_normal_and_real_psf():
    # without Toeplitz the normal operator is adjoint(forward); with the
    # real PSF the K x K PSF keeps its real parts (oracle PSF made real)
    n, C = 32, 1
    traj, B, w = case((n, n, 1), T=3, K=3, spf=8, R=64, seed=8)
    T, K = B.shape
    ref = O.BasisNufft((n, n, 1), traj, B, C)
    x = crandn((K, C, 1, n, n), 9)
    op = make_op((n, n, 1), traj, B, C, toeplitz=False)
    e_x = relerr(op.normal(dev(x)).cpu().numpy(), ref.adjoint(ref.forward(x)))
    op.free()
    op = make_op((n, n, 1), traj, B, C, real=True)
    ref.psfB = np.ascontiguousarray(ref.psfB.real.astype(np.complex64))
    e_r = relerr(op.normal(dev(x)).cpu().numpy(), ref.normal(x))
    print(f"explicit normal {e_x:.1e}, real PSF normal {e_r:.1e}")
    assert e_x < 1e-5 and e_r < 1e-5
    op.free()


def test_basis_sense_and_cgsense():
    import bart_amd as ba
    n, C = 16, 3
    traj, B, w = case((n, n, 1), T=3, K=2, spf=6, R=32, seed=11, weights=True)
    T, K = B.shape
    Mf = traj.shape[1]
    maps = I.gaussian_coils((1, n, n), C, seed=3)
    ref = O.BasisNufft((n, n, 1), traj, B, C, weights=w)
    op = ba.sense_nc_init(d16(n, n, 1, C, 1, 1, K), d16(n, n, 1, C), maps, d16(1, Mf, 1, C, 1, T),
                          d16(3, Mf, 1, 1, 1, T), np.ascontiguousarray(traj), wgs_dims=d16(1, Mf, 1, 1, 1, T),
                          weights=w, basis_dims=d16(1, 1, 1, 1, 1, T, K), basis=np.ascontiguousarray(B.T))
    x = crandn((K, 1, n, n), 12)
    y = op.forward(dev(x)).cpu().numpy()
    e_f = relerr(y, ref.sense_forward(maps, x))
    yk = crandn((T, C, Mf), 13)
    e_a = relerr(op.adjoint(dev(yk)).cpu().numpy(), ref.sense_adjoint(maps, yk))
    e_n = relerr(op.normal(dev(x)).cpu().numpy(), ref.sense_normal(maps, x))
    print(f"SENSE forward {e_f:.1e} adjoint {e_a:.1e} normal {e_n:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5 and e_n < 1e-5
    # CG-SENSE to 1e-6 against the dense solve of the oracle's normal equations
    lam = 0.05
    Nx = K * n * n
    A = np.zeros((Nx, Nx), np.complex128)
    for j in range(Nx):
        e = np.zeros(Nx, np.complex64)
        e[j] = 1
        A[:, j] = ref.sense_normal(maps, e.reshape(K, 1, n, n)).ravel()
    b = op.adjoint(dev(yk))
    xs = np.linalg.solve(A + lam * np.eye(Nx), b.cpu().numpy().ravel().astype(np.complex128))
    xg, st, _ = ba.lsqr_conjgrad(op, b, 300, lam, 1e-6)
    e_cg = relerr(xg.cpu().numpy(), xs)
    print(f"CG-SENSE {st.iterations} iterations, vs dense solve {e_cg:.1e}")
    assert st.status == 2 and e_cg < 1e-4
    # repeated solves are bitwise equal
    xg2, _, _ = ba.lsqr_conjgrad(op, b, 300, lam, 1e-6)
    assert torch.equal(xg, xg2)
    op.free()

