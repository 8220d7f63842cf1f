"""GPU: ESPIRiT (extract_calib + calib of the C ABI, bart ecalib) against the
oracle's float64 restatement (oracle/espirit.py): the number of kernels and
the singular values, the maps and eigenvalue maps away from the crop
boundary (coil counts 4 to 24: every lane-group width of the point-wise
iteration) (the crop compares an eigenvalue with 0.8 and flips where FP32
and FP64 land on either side), 2-D and 3-D, one and two maps, with and
without the phase rotation; the maps explain the coil images as the
reference's test requires (tests/ecalib.mk:3-9); the command-line tool
gives the library's result (to 1e-5: the eigensolver's reductions are not
bitwise repeatable)."""
import numpy as np
import pytest

from oracle import espirit as E
from bart_amd import inputs as I
from cli_util import dims16, read_raw, run_bart, write_raw

pytestmark = pytest.mark.gpu


def fftc(x, axes):
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x, axes=axes), axes=axes, norm="ortho"), axes=axes)


def case(dims, C, seed):
    X, Y, Z = dims
    if Z == 1:
        n = max(X, Y)
        img = I.ellipse_phantom(n, seed=seed)[:, (n - Y) // 2:(n - Y) // 2 + Y, (n - X) // 2:(n - X) // 2 + X]
    else:
        img = I.ellipsoid_phantom(X, nell=8, seed=seed)
    S = I.gaussian_coils((Z, Y, X), C, seed=seed + 1)
    cim = S * img[None]
    ksp = fftc(cim, (1, 2, 3) if Z > 1 else (2, 3)).astype(np.complex64)
    return img, S, ksp


def compare(sens, ev, rs, re, margin=0.05, crop=0.8):
    # points whose eigenvalue is clear of the crop threshold in both
    keep = (np.abs(re - crop) > margin) & (np.abs(ev - crop) > margin)
    a = sens * keep[:, None]
    b = rs * keep[:, None]
    e_m = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    e_e = float(np.linalg.norm((ev - re) * keep) / np.linalg.norm(re * keep))
    return e_m, e_e, float(keep.mean())


@pytest.mark.parametrize("dims,C,maps,rot", [((64, 64, 1), 8, 1, True), ((64, 64, 1), 8, 2, True),
                                              ((96, 80, 1), 12, 2, False), ((32, 32, 32), 6, 1, True),
                                              ((64, 64, 1), 4, 2, True), ((64, 64, 1), 24, 3, True)])
def test_ecalib_vs_oracle(dims, C, maps, rot):
    import bart_amd as ba
    img, S, ksp = case(dims, C, 7)
    conf = ba.EcalibConf.defaults(rotphase=rot)
    sens, ev, sv = ba.ecalib(ksp, maps=maps, conf=conf)
    rs, re, nk = E.ecalib(ksp, maps=maps, rotphase=rot)
    kd = [6 if n > 1 else 1 for n in dims]
    N = int(np.prod(kd)) * C
    G = E.casorati_gram(E.calib_geom(ksp.astype(np.complex128), (24, 24, 24)), kd)
    _, val, nk_ref = E.compute_kernels(G, 0.001)
    e_sv = float(np.max(np.abs(sv[:nk_ref] - val[:nk_ref]) / val[0]))
    e_m, e_e, frac = compare(sens, ev, rs, re)
    print(f"{dims} C {C} maps {maps} rot {rot}: kernels {nk_ref}, sv {e_sv:.1e}, maps {e_m:.1e}, "
          f"eigenvalues {e_e:.1e} over {frac:.2f} of the points")
    assert sv.shape == (N,) and e_sv < 1e-5
    assert e_m < 1e-4 and e_e < 1e-4
    # the first map explains the coil images (tests/ecalib.mk:3-9)
    m = sens[0]
    cim = S * img[None]
    P = np.sum(np.conj(m) * cim, axis=0)[None] * m
    assert np.linalg.norm(P - cim) / np.linalg.norm(cim) < 0.05


def test_cli_ecalib(tmp_path):
    import bart_amd as ba
    img, S, ksp = case((64, 64, 1), 8, 11)
    write_raw(str(tmp_path / "k"), dims16(64, 64, 1, 8), ksp)
    run_bart(["ecalib", "-m1", "k", "s", "e"], tmp_path)
    ds, s = read_raw(str(tmp_path / "s"))
    de, e = read_raw(str(tmp_path / "e"))
    assert ds == dims16(64, 64, 1, 8, 1) and de == dims16(64, 64, 1, 1, 1)
    sens, ev, _ = ba.ecalib(ksp, maps=1)
    sens2, ev2b, _ = ba.ecalib(ksp, maps=1)
    print("repeat: bitwise equal", np.array_equal(sens, sens2),
          "max diff", float(np.abs(sens - sens2).max()))
    e_cli, e_ev, _ = compare(s.reshape(sens.shape), e.real.reshape(ev.shape), sens, ev)
    print(f"CLI vs library: maps {e_cli:.1e}, eigenvalues {e_ev:.1e}")
    assert e_cli < 1e-5 and e_ev < 1e-5
    # -P: no phase rotation; -r / -k / -t / -c as in the reference
    run_bart(["ecalib", "-P", "-m1", "-r", "20:20:20", "-k", "5:5:5", "-t", "0.002", "-c", "0.7", "k", "s2"],
             tmp_path)
    _, s2 = read_raw(str(tmp_path / "s2"))
    rs, re, _ = E.ecalib(ksp, maps=1, calsize=(20, 20, 20), kdims=(5, 5, 5), threshold=0.002, crop=0.7,
                         rotphase=False)
    sg = s2.reshape(1, 8, 1, 64, 64)
    ev2 = ba.ecalib(ksp, maps=1, calsize=(20, 20, 20),
                    conf=ba.EcalibConf.defaults(rotphase=False, kdims=(5, 5, 5), threshold=0.002, crop=0.7))[1]
    e_m, _, _ = compare(sg, ev2, rs, re, crop=0.7)
    assert e_m < 1e-4
    # unsupported options are errors
    p = run_bart(["ecalib", "-W", "k", "s3"], tmp_path, check=False)
    assert p.returncode == 1 and "-W" in p.stderr
