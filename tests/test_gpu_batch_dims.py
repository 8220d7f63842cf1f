"""GPU: batch dims >= 4 that share one trajectory (several slices or frames
in one operator, as BART's NUFFT treats them like coils and pics runs one
CG over the whole batch): NUFFT forward, adjoint and Toeplitz normal
against the oracle with C x B channels; the SENSE operator (maps applied
to every image) against the oracle's per-image SENSE operators; CG-SENSE
over the batch against per-image dense solves; the reference's pics
low-memory test through the command line."""
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


@pytest.mark.parametrize("dims", [(32, 32, 1), (16, 16, 16)])
def test_nufft_batch_dims_vs_oracle(dims):
    import bart_amd as ba
    C, B5, B6 = 2, 3, 2
    traj = I.phyllotaxis_kooshball(spokes=60, readout=32) if dims[2] > 1 else I.radial_traj(64, 21)
    M = traj.shape[0]
    w = (0.5 + np.random.default_rng(1).random(M)).astype(np.complex64)
    op = ba.nufft_create2(d16(1, M, 1, C, 1, B5, B6), d16(*dims, C, 1, B5, B6), d16(3, M, 1),
                          np.ascontiguousarray(traj), wgh_dims=d16(1, M, 1), weights=w)
    ref = O.Nufft(dims, traj, C * B5 * B6, weights=w, toeplitz=True)
    x = crandn((C * B5 * B6,) + dims[::-1], 2)
    y = crandn((C * B5 * B6, M), 3)
    e_f = relerr(op.forward(dev(x)).cpu().numpy(), ref.forward(x))
    e_a = relerr(op.adjoint(dev(y)).cpu().numpy(), ref.adjoint(y))
    e_n = relerr(op.normal(dev(x)).cpu().numpy(), ref.normal(x))
    print(f"{dims} batch 3 x 2: forward {e_f:.1e} adjoint {e_a:.1e} normal {e_n:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5 and e_n < 1e-5
    op.free()


def test_sense_batch_dims_and_cgsense():
    import bart_amd as ba
    n, C, B = 16, 3, 4
    traj = I.radial_traj(32, 13)
    M = traj.shape[0]
    maps = I.gaussian_coils((1, n, n), C, seed=4)
    op = ba.sense_nc_init(d16(n, n, 1, C, 1, B), d16(n, n, 1, C), maps, d16(1, M, 1, C, 1, B), d16(3, M, 1),
                          np.ascontiguousarray(traj))
    ref = O.Nufft((n, n, 1), traj, C, toeplitz=True)
    x = crandn((B, 1, n, n), 5)
    y = crandn((B, C, M), 6)
    f = op.forward(dev(x)).cpu().numpy().reshape(B, C, M)
    a = op.adjoint(dev(y)).cpu().numpy().reshape(B, n, n)
    nx = op.normal(dev(x)).cpu().numpy().reshape(B, n, n)
    e_f = relerr(f, np.stack([ref.sense_forward(maps, x[b]) for b in range(B)]))
    e_a = relerr(a, np.stack([ref.sense_adjoint(maps, y[b]) for b in range(B)]))
    e_n = relerr(nx, np.stack([ref.sense_normal(maps, x[b]) for b in range(B)]))
    print(f"SENSE batch {B}: forward {e_f:.1e} adjoint {e_a:.1e} normal {e_n:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5 and e_n < 1e-5
    # one CG over the batch (as pics runs it) reaches the per-image solutions
    lam = 0.05
    A = np.zeros((n * n, n * n), np.complex128)
    for j in range(n * n):
        e = np.zeros(n * n, np.complex64)
        e[j] = 1
        A[:, j] = ref.sense_normal(maps, e.reshape(1, n, n)).ravel()
    b = op.adjoint(dev(y))
    bh = b.cpu().numpy().reshape(B, n * n).astype(np.complex128)
    xs = np.stack([np.linalg.solve(A + lam * np.eye(n * n), bh[i]) for i in range(B)])
    xg, st, _ = ba.lsqr_conjgrad(op, b, 300, lam, 1e-6)
    e_cg = relerr(xg.cpu().numpy(), xs)
    print(f"CG-SENSE over the batch: {st.iterations} iterations, vs dense solves {e_cg:.1e}")
    assert st.status == 2 and e_cg < 1e-4
    op.free()


def test_cli_pics_batch_lowmem(tmp_path):
    # tests/pics.mk:330-347 (test-pics-lowmem) at a smaller batch: k-space
    # repeated over dims 5 and 6, -U gives the same reconstruction
    run_bart(["traj", "-x", "16", "-y", "3", "-D", "-r", "t"], tmp_path)
    td, t = read_raw(str(tmp_path / "t"))
    n = 8
    write_raw(str(tmp_path / "x"), dims16(n, n, 1), I.ellipse_phantom(n, seed=3))
    run_bart(["nufft", "-x", f"{n}:{n}:1", "t", "x", "k0"], tmp_path)
    kd, k0 = read_raw(str(tmp_path / "k0"))
    B5, B6 = 20, 20
    write_raw(str(tmp_path / "k2"), kd[:5] + [B5, B6] + [1] * 9, np.tile(k0, B5 * B6))
    write_raw(str(tmp_path / "s"), dims16(n, n), np.ones(n * n, np.complex64))
    run_bart(["pics", "-U", "-t", "t", "k2", "s", "r1"], tmp_path)
    run_bart(["pics", "-t", "t", "k2", "s", "r2"], tmp_path)
    d1, r1 = read_raw(str(tmp_path / "r1"))
    _, r2 = read_raw(str(tmp_path / "r2"))
    assert d1 == dims16(n, n, 1, 1, 1, B5, B6)
    e = relerr(r1, r2)
    # every image of the batch is the reconstruction of the same data
    rr = r2.reshape(B5 * B6, n * n)
    spread = max(relerr(rr[i], rr[0]) for i in range(B5 * B6))
    print(f"pics -U vs pics over a {B5} x {B6} batch: {e:.1e}; spread over the batch {spread:.1e}")
    assert e < 5e-4 and spread < 1e-6


def frame_trajs(F, R, S):
    # golden-angle radial, a different set of spokes per frame
    t = I.radial_traj(R, F * S)
    return np.ascontiguousarray(t.reshape(F, R * S, 3))


def test_nufft_frame_trajectories_vs_oracle():
    # a trajectory per frame (dims >= 4 of the trajectory equal to the
    # k-space's): frame t is gridded with its own trajectory and weights
    import bart_amd as ba
    n, C, F, R, S = 32, 2, 3, 64, 7
    traj = frame_trajs(F, R, S)
    M = R * S
    w = (0.5 + np.random.default_rng(7).random((F, M))).astype(np.complex64)
    op = ba.nufft_create2(d16(1, R, S, C, 1, F), d16(n, n, 1, C, 1, F), d16(3, R, S, 1, 1, F), traj,
                          wgh_dims=d16(1, R, S, 1, 1, F), weights=w)
    x = crandn((F, C, 1, n, n), 8)
    y = crandn((F, C, M), 9)
    f = op.forward(dev(x)).cpu().numpy().reshape(F, C, M)
    a = op.adjoint(dev(y)).cpu().numpy().reshape(F, C, 1, n, n)
    nx = op.normal(dev(x)).cpu().numpy().reshape(F, C, 1, n, n)
    refs = [O.Nufft((n, n, 1), traj[t], C, weights=w[t], toeplitz=True) for t in range(F)]
    e_f = relerr(f, np.stack([refs[t].forward(x[t]) for t in range(F)]))
    e_a = relerr(a, np.stack([refs[t].adjoint(y[t]) for t in range(F)]))
    e_n = relerr(nx, np.stack([refs[t].normal(x[t]) for t in range(F)]))
    print(f"frames: forward {e_f:.1e} adjoint {e_a:.1e} normal {e_n:.1e}")
    assert e_f < 1e-5 and e_a < 1e-5 and e_n < 1e-5
    op.free()


def test_sense_frame_trajectories_cgsense():
    import bart_amd as ba
    n, C, F, R, S = 16, 3, 3, 32, 9
    traj = frame_trajs(F, R, S)
    M = R * S
    maps = I.gaussian_coils((1, n, n), C, seed=8)
    op = ba.sense_nc_init(d16(n, n, 1, C, 1, F), d16(n, n, 1, C), maps, d16(1, R, S, C, 1, F),
                          d16(3, R, S, 1, 1, F), traj)
    refs = [O.Nufft((n, n, 1), traj[t], C, toeplitz=True) for t in range(F)]
    x = crandn((F, 1, n, n), 10)
    y = crandn((F, C, M), 11)
    e_n = relerr(op.normal(dev(x)).cpu().numpy().reshape(F, n, n),
                 np.stack([refs[t].sense_normal(maps, x[t])[0] for t in range(F)]))
    assert e_n < 1e-5
    lam = 0.05
    b = op.adjoint(dev(y))
    bh = b.cpu().numpy().reshape(F, n * n).astype(np.complex128)
    xs = []
    for t in range(F):
        A = np.zeros((n * n, n * n), np.complex128)
        for j in range(n * n):
            e = np.zeros(n * n, np.complex64)
            e[j] = 1
            A[:, j] = refs[t].sense_normal(maps, e.reshape(1, n, n)).ravel()
        xs.append(np.linalg.solve(A + lam * np.eye(n * n), bh[t]))
    xg, st, _ = ba.lsqr_conjgrad(op, b, 300, lam, 1e-6)
    e_cg = relerr(xg.cpu().numpy(), np.stack(xs))
    print(f"frames SENSE normal {e_n:.1e}; CG {st.iterations} iterations, vs per-frame dense solves {e_cg:.1e}")
    assert st.status == 2 and e_cg < 1e-4
    op.free()
