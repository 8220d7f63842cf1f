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
    (["nufft", "-B", "b", "t", "i", "o"], "-B"),
    (["pics", "-l1", "-t", "t", "k", "s", "o"], "-l1"),
    (["pics", "-R", "W:7:0:0.01", "-t", "t", "k", "s", "o"], "-R"),
    (["pics", "k", "s", "o"], "non-Cartesian"),
    (["pics", "--admm", "-t", "t", "k", "s", "o"], "admm"),
    (["nosuchtool"], "Unknown bart command"),
])
def test_cli_rejects_unsupported(tmp_path, args, msg):
    p = run_bart(args, tmp_path, check=False)
    assert p.returncode == 1
    assert msg in p.stderr


def test_cli_missing_file_is_an_error(tmp_path):
    p = run_bart(["nufft", "-a", "nope", "nope", "out"], tmp_path, check=False)
    assert p.returncode == 1 and "ERROR" in p.stderr
