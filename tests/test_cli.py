"""Command-line drop-in on the CPU: trajectory generation (no GPU needed),
BART file formats, and option handling that fails before any GPU work."""
import numpy as np
import pytest

from bart_amd import cfl
from bart_amd import inputs as I
from cli_util import read_raw, run_bart, write_raw


def test_traj_golden_radial_matches_formula(tmp_path):
    # traj -r -G -x 64 -o 2 -y 21 (src/traj.c:303-455)
    run_bart(["traj", "-r", "-G", "-x", 64, "-o", 2, "-y", 21, "t"], tmp_path)
    dims, t = read_raw(str(tmp_path / "t"))
    assert dims[:3] == [3, 128, 21] and all(d == 1 for d in dims[3:])
    ref = I.radial_traj(128, 21).reshape(-1)
    # sinf/cosf of the C tool vs numpy's float32 sin/cos: within 1 ulp
    assert np.max(np.abs(t - ref)) <= 2e-6 * np.max(np.abs(ref))
    assert np.all(t.imag == 0)


def test_traj_uniform_radial_and_cartesian(tmp_path):
    run_bart(["traj", "-r", "-x", 32, "-y", 16, "u"], tmp_path)
    _, t = read_raw(str(tmp_path / "u"))
    ref = I.radial_traj(32, 16, over=1.0, golden=False).reshape(-1)
    assert np.max(np.abs(t - ref)) <= 2e-6 * np.max(np.abs(ref))
    run_bart(["traj", "-x", 8, "-y", 6, "c"], tmp_path)
    dims, t = read_raw(str(tmp_path / "c"))
    assert dims[:3] == [3, 8, 6]
    assert np.array_equal(t, I.cartesian_traj(8, 6).reshape(-1))


def test_traj_kooshball_points_on_sphere(tmp_path):
    # traj -3 -r: Saff-Kuijlaars spoke directions; every spoke is a line
    # through the centre with |k| = |read| / over
    run_bart(["traj", "-3", "-r", "-x", 16, "-y", 40, "k"], tmp_path)
    dims, t = read_raw(str(tmp_path / "k"))
    k = t.real.reshape(40, 16, 3)
    read = np.arange(16) + 0.5 - 8
    r = np.linalg.norm(k, axis=-1)
    assert np.allclose(r, np.abs(read)[None, :], rtol=1e-5)
    z = k[:, -1, 2] / read[-1]
    # elevation acos(hn) - pi/2 with hn from -1 (first spoke, at the pole)
    # towards 0: z falls from 1 to about 0 over the half sphere
    assert np.isclose(z[0], 1.0, atol=1e-6) and np.all(np.diff(z) < 0) and z[-1] > 0


def test_traj_invariants_of_the_reference_tests(tmp_path):
    # tests/traj.mk test-traj-over: -x128 -r scaled by 0.5 equals -x64 -o2 -r
    run_bart(["traj", "-x", 128, "-y", 71, "-r", "a"], tmp_path)
    run_bart(["traj", "-x", 64, "-y", 71, "-o", 2.0, "-r", "b"], tmp_path)
    _, a = read_raw(str(tmp_path / "a"))
    _, b = read_raw(str(tmp_path / "b"))
    assert np.linalg.norm(0.5 * a - b) <= 1e-7 * np.linalg.norm(b)
    # test-traj-3D: the first sample of every kooshball spoke lies at radius 63.5
    run_bart(["traj", "-3", "-x", 128, "-y", 128, "-r", "k"], tmp_path)
    _, k = read_raw(str(tmp_path / "k"))
    k = k.real.reshape(128, 128, 3)
    r0 = np.linalg.norm(k[:, 0, :].astype(np.float64), axis=-1)
    assert np.linalg.norm(r0 - 63.5) <= 1e-7 * np.linalg.norm(np.full(128, 63.5))


def test_traj_golden_angle_increment(tmp_path):
    # -G: consecutive spokes turn by the golden angle pi / phi (phi the golden
    # ratio, 111.246 degrees; modulo pi, as a spoke through the centre has no
    # direction), -H by the half-circle golden angle; -r alone: pi / spokes
    phi = (1 + np.sqrt(5.0)) / 2
    for flags, inc in ((["-G"], np.pi / phi), ([], np.pi / 21)):
        run_bart(["traj", "-r", *flags, "-x", 64, "-y", 21, "g"], tmp_path)
        _, t = read_raw(str(tmp_path / "g"))
        k = t.real.reshape(21, 64, 3).astype(np.float64)
        ang = np.arctan2(k[:, -1, 1], k[:, -1, 0])
        # (the turning direction is BART's: negative)
        d = np.mod(np.diff(ang) + inc, np.pi)
        d = np.minimum(d, np.pi - d)
        assert np.max(d) < 1e-5, (flags, np.max(d))


def test_cfl_and_ra_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((5, 4, 3)) + 1j * rng.standard_normal((5, 4, 3))).astype(np.complex64)
    cfl.writecfl(str(tmp_path / "a"), a)
    cfl.writecfl(str(tmp_path / "a.ra"), a)
    b = cfl.readcfl(str(tmp_path / "a"))
    c = cfl.readcfl(str(tmp_path / "a.ra"))
    assert b.shape == (5, 4, 3) and np.array_equal(a, b) and np.array_equal(a, c)
    # the C tools read what Python wrote and write the same bytes as .ra
    run_bart(["traj", "-x", 4, "-y", 3, "t.ra"], tmp_path)
    run_bart(["traj", "-x", 4, "-y", 3, "t"], tmp_path)
    assert np.array_equal(cfl.readcfl(str(tmp_path / "t.ra")), cfl.readcfl(str(tmp_path / "t")))
    with open(tmp_path / "t.ra", "rb") as f:
        head = np.frombuffer(f.read(48), dtype="<u8")
    assert head[0] == 0x7961727261776172 and head[2] == 4 and head[3] == 8 and head[4] == 8 * 36


def test_header_with_extra_sections(tmp_path):
    # BART writes "# Command", "# Files", "# Creator" sections after the dims
    p = str(tmp_path / "h")
    write_raw(p, [3, 2, 1], np.arange(6, dtype=np.complex64))
    with open(p + ".hdr", "a") as f:
        f.write("# Command\nbart traj -x 2 h\n# Creator\nBART v0.9\n")
    assert cfl.readcfl(p).shape == (3, 2)


@pytest.mark.parametrize("args,msg", [
    (["nufft", "-s", "-p", "p", "t", "i", "o"], "ERROR"),
    (["nufft", "-a", "-i", "t", "i", "o"], "Adjoint and inverse"),
    (["nufft", "-F", "f", "t", "i", "o"], "-F"),
    (["pics", "-l1", "-t", "t", "k", "s", "o"], "-l1"),
    (["pics", "-R", "W:7:0:0.01", "-t", "t", "k", "s", "o"], "-R"),
    (["pics", "k", "s", "o"], "non-Cartesian"),
    (["pics", "--admm", "-t", "t", "k", "s", "o"], "admm"),
    (["nosuchtool"], "Unknown bart command"),
    (["nufft", "--nufft-conf", "bogus", "t", "i", "o"], "Sub-option could not be parsed: bogus"),
])
def test_cli_rejects_unsupported(tmp_path, args, msg):
    p = run_bart(args, tmp_path, check=False)
    assert p.returncode == 1
    assert msg in p.stderr


def test_cli_missing_file_is_an_error(tmp_path):
    p = run_bart(["nufft", "-a", "nope", "nope", "out"], tmp_path, check=False)
    assert p.returncode == 1 and "ERROR" in p.stderr
