"""GPU: batched CG-SENSE (lsqr_conjgrad_batch) over problems of uneven image
size, coil count and spoke count.

Each problem of a batch must follow exactly the iterations of its single
solve (lsqr_conjgrad): same solution bits, iteration count, stop status and
residual trace.  The single solve itself is checked against the CPU oracle in
test_gpu_nufft.py; here one batched problem is also compared with the oracle
directly, at the tolerance stated there (6 iterations, 1e-4).
"""
import numpy as np
import pytest

import bart_amd as ba
import oracle as O
from bart_amd import inputs as I
from conftest import crandn, relerr

pytestmark = pytest.mark.gpu


def dims16(*d):
    return list(d) + [1] * (16 - len(d))


def sense_problem(n, C, R, S, Z=1, seed=0, noise=1e-3):
    if Z == 1:
        traj = I.radial_traj(R, S)
        img = I.ellipse_phantom(n, seed=seed)
    else:
        traj = I.phyllotaxis_kooshball(spokes=S, readout=R)
        img = I.ellipsoid_phantom(n, nell=6, seed=seed)
    maps = I.gaussian_coils((Z, n, n), C, seed=seed)
    op = ba.sense_nc_init(dims16(n, n, Z, C), dims16(n, n, Z, C), maps, dims16(1, R, S, C), dims16(3, R, S),
                          ba.traj_array(traj))
    y = op.forward(img).reshape(C, R * S)
    y = I.add_noise(y, noise, seed=seed)
    b = op.adjoint(y)
    return dict(op=op, b=b, n=n, Z=Z, C=C, traj=traj, maps=maps)


# (n, C, readout, spokes, Z): uneven sizes, coils and spokes; one 3-D problem
SHAPES = [(32, 2, 64, 21, 1), (48, 4, 96, 31, 1), (64, 3, 128, 41, 1), (40, 8, 80, 25, 1),
          (64, 2, 128, 61, 1), (32, 5, 64, 11, 1), (16, 2, 32, 60, 16)]


@pytest.fixture(scope="module")
def problems():
    ps = [sense_problem(n, C, R, S, Z, seed=100 + i) for i, (n, C, R, S, Z) in enumerate(SHAPES)]
    yield ps
    for p in ps:
        p["op"].free()


def same(a, b):
    return np.array_equal(np.asarray(a).reshape(-1), np.asarray(b).reshape(-1))


@pytest.mark.parametrize("tol,maxiter", [(0.0, 12), (1e-4, 150)])
def test_batch_equals_single_solves(problems, tol, maxiter):
    ops = [p["op"] for p in problems]
    bs = [p["b"] for p in problems]
    lam = 0.01
    xs, st, tr = ba.lsqr_conjgrad_batch(ops, bs, maxiter, lam, tol)
    its = []
    for i, p in enumerate(problems):
        x1, s1, t1 = ba.lsqr_conjgrad(p["op"], p["b"], maxiter, lam, tol)
        assert st[i].iterations == s1.iterations, i
        assert st[i].status == s1.status, i
        assert same(xs[i], x1), i
        assert same(tr[i], t1), i
        its.append(s1.iterations)
    if tol > 0:
        # the problems stop at different iterations, each by its own test
        assert sum(s.status == 2 for s in st) >= 2
        assert len(set(its)) > 1
        for s in st:
            if s.status == 2:
                assert s.resid <= 1e-4 * s.bnorm * 1.0001


def test_batch_problem_matches_oracle(problems):
    p = problems[2]
    n, C = p["n"], p["C"]
    ref = O.Nufft((n, n, 1), p["traj"], C, toeplitz=True)
    b = np.asarray(p["b"]).reshape(1, n, n)
    xr6, _, _, _ = ref.cgsense(p["maps"], b, 0.01, 6, 0.0)
    xs, st, _ = ba.lsqr_conjgrad_batch([q["op"] for q in problems], [q["b"] for q in problems], 6, 0.01, 0.0)
    assert st[2].iterations == 6
    assert relerr(xs[2].reshape(1, n, n), xr6) < 1e-4


def test_batch_warm_start_and_fallback(problems):
    ps = problems[:4]
    ops = [p["op"] for p in ps]
    bs = [p["b"] for p in ps]
    lam, tol = 0.01, 1e-4
    sol = [ba.lsqr_conjgrad(p["op"], p["b"], 60, lam, tol)[0] for p in ps]
    # warm starts: the solution, a bad guess (falls back to zero), zero, a
    # perturbed solution
    x0 = [sol[0], (1e3 * crandn(np.shape(sol[1]), 7)).astype(np.complex64), np.zeros_like(sol[2]),
          (sol[3] + 0.05 * crandn(np.shape(sol[3]), 8)).astype(np.complex64)]
    xs, st, tr = ba.lsqr_conjgrad_batch(ops, bs, 60, lam, tol, x0s=x0, fallback=True)
    for i, p in enumerate(ps):
        x1, s1, t1 = ba.lsqr_conjgrad(p["op"], p["b"], 60, lam, tol, x0=x0[i], fallback=True)
        assert st[i].fell_back == s1.fell_back, i
        assert st[i].iterations == s1.iterations and st[i].status == s1.status, i
        assert same(xs[i], x1) and same(tr[i], t1), i
    assert st[1].fell_back and not st[0].fell_back
    assert st[0].iterations <= 2
    # 0 iterations return the warm starts unchanged (tests/pics.mk:184-190)
    xz, stz, _ = ba.lsqr_conjgrad_batch(ops, bs, 0, lam, 0.0, x0s=x0)
    for i in range(len(ps)):
        assert stz[i].iterations == 0 and same(xz[i], x0[i])


def test_batch_zero_rhs_and_single_problem(problems):
    p0, p1 = problems[0], problems[1]
    zero = np.zeros_like(np.asarray(p0["b"]))
    xs, st, tr = ba.lsqr_conjgrad_batch([p0["op"], p1["op"]], [zero, p1["b"]], 10, 0.01, 0.0)
    assert st[0].status == 4 and st[0].iterations == 0 and not np.any(xs[0])
    assert st[1].iterations == 10
    x1, s1, _ = ba.lsqr_conjgrad(p1["op"], p1["b"], 10, 0.01, 0.0)
    assert same(xs[1], x1)
    # a batch of one and an empty batch
    xo, so, _ = ba.lsqr_conjgrad_batch([p1["op"]], [p1["b"]], 10, 0.01, 0.0)
    assert same(xo[0], x1)
    xe, se, _ = ba.lsqr_conjgrad_batch([], [], 10, 0.01, 0.0)
    assert xe == [] and se == []


def test_batch_rejects_shared_plan(problems):
    p = problems[0]
    with pytest.raises(ba.BartError, match="share"):
        ba.lsqr_conjgrad_batch([p["op"], p["op"]], [p["b"], p["b"]], 5, 0.01, 0.0)


def test_batch_device_vectors(problems):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU for torch")
    ps = problems[:3]
    bd = [torch.from_numpy(np.asarray(p["b"]).reshape(-1)).to("cuda") for p in ps]
    xs, st, _ = ba.lsqr_conjgrad_batch([p["op"] for p in ps], bd, 8, 0.01, 0.0)
    torch.cuda.synchronize()
    for i, p in enumerate(ps):
        x1, _, _ = ba.lsqr_conjgrad(p["op"], p["b"], 8, 0.01, 0.0)
        assert same(xs[i].cpu().numpy(), x1)


def test_batch_mixed_normal_operators(problems):
    # a SENSE operator without Toeplitz PSF (normal = adjoint(forward(x)),
    # linops/linop.c:489-525) beside Toeplitz ones: each problem still follows
    # its single solve bit for bit
    n, C, R, S = 48, 3, 96, 27
    traj = I.radial_traj(R, S)
    maps = I.gaussian_coils((1, n, n), C, seed=77)
    conf = ba.NufftConf.defaults(toeplitz=False)
    op = ba.sense_nc_init(dims16(n, n, 1, C), dims16(n, n, 1, C), maps, dims16(1, R, S, C), dims16(3, R, S),
                          ba.traj_array(traj), conf=conf)
    y = I.add_noise(op.forward(I.ellipse_phantom(n, seed=77)).reshape(C, R * S), 1e-3, seed=77)
    b = op.adjoint(y)
    ops = [problems[0]["op"], op, problems[2]["op"]]
    bs = [problems[0]["b"], b, problems[2]["b"]]
    xs, st, tr = ba.lsqr_conjgrad_batch(ops, bs, 15, 0.01, 1e-4)
    for i, (o, bb) in enumerate(zip(ops, bs)):
        x1, s1, t1 = ba.lsqr_conjgrad(o, bb, 15, 0.01, 1e-4)
        assert st[i].iterations == s1.iterations and same(xs[i], x1) and same(tr[i], t1), i
    op.free()


def test_adjoint_gridding_autotuner_report():
    # every plan's adjoint gridding launch is tuned when the plan is built
    # (coils per work-group 4 .. default, work-groups per CU 1, 2, 4); the
    # choice is reported per shape class and the adjoint keeps its accuracy
    n, C, R, S = 96, 6, 192, 33
    traj = I.radial_traj(R, S)
    op = ba.nufft_create2(dims16(1, R, S, C), dims16(n, n, 1, C), dims16(3, R, S), ba.traj_array(traj),
                          conf=ba.NufftConf.defaults(toeplitz=False))
    rep = ba.tuning_report()
    key = f"adj:{n}x{n}x1:S4:C{C}:M{R * S}"
    mine = [e for e in rep if e["shape"].startswith(key)]
    assert mine, rep
    e = mine[-1]
    assert e["coils_per_workgroup"] in (4, 8) and e["workgroups_per_cu"] in (1, 2, 4)
    y = crandn((C, R * S), 71)
    ref = O.Nufft((n, n, 1), traj, C, toeplitz=False)
    assert relerr(op.adjoint(y).reshape(C, 1, n, n), ref.adjoint(y)) < 1e-5
    op.free()
