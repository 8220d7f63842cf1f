"""Seeded synthetic inputs for the NUFFT / CG-SENSE workloads.

BART's own `phantom` and `traj` tools do not generate these inputs (random
ellipse phantoms, Gaussian ring coils, phyllotaxis kooshballs, variable-density
spirals), so they are generated here and handed identically to the HIP path and
to the CPU oracle.  Arrays follow the C-ordered view of BART's column-major
layout: images [Z, Y, X], coil images / maps [C, Z, Y, X], k-space [C, M],
trajectory [M, 3] complex64 with the coordinate in the real part.

The radial trajectory follows BART's `traj -r -G` formula (src/traj.c:303-455:
read = i + 0.5 - R/2, read direction (sin phi, cos phi, 0), golden angle
pi (sqrt(5) - 1) / 2 from noncart/traj.c:154-181), divided by the readout
oversampling.
"""
import numpy as np

GOLDEN_ANGLE_2D = np.pi * (np.sqrt(5.0) - 1.0) / 2.0


def radial_traj(readout=512, spokes=201, over=2.0, golden=True):
    """Golden-angle (or uniform) 2-D radial trajectory, [readout * spokes, 3]."""
    i = np.arange(readout, dtype=np.float64)
    read = i + 0.5 - readout / 2.0
    j = np.arange(spokes, dtype=np.float64)
    if golden:
        ang = (j * GOLDEN_ANGLE_2D).astype(np.float32)
    else:
        ang = (j * np.pi / spokes).astype(np.float32)
    sx = np.sin(ang).astype(np.float32)
    cy = np.cos(ang).astype(np.float32)
    kx = (read[:, None] * sx[None, :]) / over
    ky = (read[:, None] * cy[None, :]) / over
    t = np.zeros((spokes, readout, 3), np.complex64)   # m = readout + R * spoke
    t[..., 0] = kx.T.astype(np.float32)
    t[..., 1] = ky.T.astype(np.float32)
    return t.reshape(-1, 3)


def cartesian_traj(nx, ny, nz=1):
    """BART `traj -x nx -y ny` Cartesian grid (src/traj.c:438-446)."""
    t = np.zeros((nz, ny, nx, 3), np.complex64)
    x = np.arange(nx) - nx // 2
    y = np.arange(ny) - ny // 2
    z = np.arange(nz) - nz // 2
    t[..., 0] = x[None, None, :]
    t[..., 1] = y[None, :, None]
    t[..., 2] = z[:, None, None]
    return t.reshape(-1, 3)


def phyllotaxis_kooshball(spokes=8192, readout=256, over=2.0):
    """3-D radial kooshball with spiral-phyllotaxis spoke ordering
    (Piccini et al., MRM 66:1049 (2011)) on the half sphere; |k| <= readout/(2 over)."""
    n = np.arange(spokes, dtype=np.float64)
    golden = (1.0 + np.sqrt(5.0)) / 2.0
    phi = 2.0 * np.pi * n / golden ** 2
    theta = 0.5 * np.pi * np.sqrt(n / spokes)
    d = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)
    read = (np.arange(readout, dtype=np.float64) + 0.5 - readout / 2.0) / over
    k = read[None, :, None] * d[:, None, :]
    t = np.zeros((spokes, readout, 3), np.complex64)
    t.real = k.astype(np.float32)
    return t.reshape(-1, 3)


def vd_spiral(interleaves=24, samples=4096, kmax=None, n=320, alpha=8.0, turns=None):
    """2-D variable-density spiral with sample density ~ 1 / (1 + alpha |k| / kmax).
    The radius follows the closed form of dr/dt ~ 1 + alpha r / kmax for a
    constant angular rate."""
    if kmax is None:
        kmax = n / 2.0 - 1.0
    if turns is None:
        turns = n / (2.0 * interleaves) * 1.5
    t = (np.arange(samples, dtype=np.float64) + 0.5) / samples
    # r' = c (1 + alpha r / kmax)  =>  r = kmax / alpha * (exp(alpha c t / kmax) - 1)
    c = kmax / alpha * np.log1p(alpha) / 1.0
    r = kmax / alpha * (np.exp(alpha * c * t / kmax) - 1.0)
    out = np.zeros((interleaves, samples, 3), np.complex64)
    for l in range(interleaves):
        th = 2.0 * np.pi * turns * t + 2.0 * np.pi * l / interleaves
        out[l, :, 0] = (r * np.cos(th)).astype(np.float32)
        out[l, :, 1] = (r * np.sin(th)).astype(np.float32)
    return out.reshape(-1, 3)


def _smooth_phase(rng, shape, amp=np.pi):
    coords = np.meshgrid(*[np.linspace(-1, 1, s) for s in shape], indexing="ij")
    ph = np.zeros(shape)
    for ax in range(len(shape)):
        ph += rng.uniform(-1, 1) * coords[ax]
        ph += 0.5 * rng.uniform(-1, 1) * coords[ax] ** 2
    m = np.max(np.abs(ph))
    if m > 0:
        ph *= amp * rng.uniform(0.3, 1.0) / m
    return np.exp(1j * ph)


def ellipse_phantom(n, nell=10, seed=0, phase=True):
    """n x n complex phantom of `nell` random ellipses (centres U[-0.6,0.6],
    semi-axes U[0.05,0.4], angle U[0,pi), intensity U[0,1], FOV = [-1,1]^2)
    times a smooth random phase.  Returns [1, n, n] (Z, Y, X)."""
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.linspace(-1, 1, n, endpoint=False) + 1.0 / n,
                       np.linspace(-1, 1, n, endpoint=False) + 1.0 / n, indexing="ij")
    img = np.zeros((n, n))
    for _ in range(nell):
        cx, cy = rng.uniform(-0.6, 0.6, 2)
        a, b = rng.uniform(0.05, 0.4, 2)
        th = rng.uniform(0, np.pi)
        inten = rng.uniform(0, 1)
        xr = (x - cx) * np.cos(th) + (y - cy) * np.sin(th)
        yr = -(x - cx) * np.sin(th) + (y - cy) * np.cos(th)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += inten
    out = img.astype(np.complex128)
    if phase:
        out = out * _smooth_phase(rng, (n, n))
    return out.astype(np.complex64)[None]


def ellipsoid_phantom(n, nell=16, seed=0, phase=True):
    """n^3 random-ellipsoid phantom, [n, n, n] (Z, Y, X)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(-1, 1, n, endpoint=False) + 1.0 / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    img = np.zeros((n, n, n), np.float32)
    for _ in range(nell):
        c = rng.uniform(-0.5, 0.5, 3)
        ax = rng.uniform(0.08, 0.4, 3)
        a1, a2 = rng.uniform(0, np.pi, 2)
        inten = rng.uniform(0, 1)
        R1 = np.array([[np.cos(a1), -np.sin(a1), 0], [np.sin(a1), np.cos(a1), 0], [0, 0, 1]])
        R2 = np.array([[1, 0, 0], [0, np.cos(a2), -np.sin(a2)], [0, np.sin(a2), np.cos(a2)]])
        R = R1 @ R2
        X = np.stack([x - c[0], y - c[1], z - c[2]], -1).astype(np.float32)
        Xr = X @ R.astype(np.float32)
        img[np.sum((Xr / ax.astype(np.float32)) ** 2, -1) <= 1.0] += inten
    out = img.astype(np.complex64)
    if phase:
        out = (out * _smooth_phase(rng, (n, n, n))).astype(np.complex64)
    return out


def gaussian_coils(shape, ncoils, seed=0, radius=0.6, sigma=0.7):
    """Gaussian coil sensitivities centred on a ring (2-D) or on a sphere
    (3-D) of the FOV [-1,1]^d, each with a smooth phase, normalised to
    root-sum-of-squares 1.  shape = (Z, Y, X).  Returns [C, Z, Y, X]."""
    rng = np.random.default_rng(seed + 7919)
    Z, Y, X = shape
    gz = np.linspace(-1, 1, Z, endpoint=False) + 1.0 / Z if Z > 1 else np.zeros(1)
    gy = np.linspace(-1, 1, Y, endpoint=False) + 1.0 / Y
    gx = np.linspace(-1, 1, X, endpoint=False) + 1.0 / X
    z, y, x = np.meshgrid(gz, gy, gx, indexing="ij")
    maps = np.zeros((ncoils, Z, Y, X), np.complex128)
    golden = (1.0 + np.sqrt(5.0)) / 2.0
    for c in range(ncoils):
        if Z == 1:
            a = 2 * np.pi * c / ncoils
            cc = (radius * np.cos(a), radius * np.sin(a), 0.0)
        else:
            h = -1.0 + (2.0 * c + 1.0) / ncoils
            a = 2 * np.pi * c / golden
            rr = np.sqrt(max(0.0, 1.0 - h * h))
            cc = (radius * rr * np.cos(a), radius * rr * np.sin(a), radius * h)
        d2 = (x - cc[0]) ** 2 + (y - cc[1]) ** 2 + (z - cc[2]) ** 2
        ph = rng.uniform(-1, 1) * x + rng.uniform(-1, 1) * y + rng.uniform(0, 2 * np.pi)
        maps[c] = np.exp(-d2 / (2 * sigma ** 2)) * np.exp(1j * ph)
    rss = np.sqrt(np.sum(np.abs(maps) ** 2, 0))
    return (maps / rss[None]).astype(np.complex64)


def add_noise(ksp, rel_sigma=1e-3, seed=0):
    """Complex Gaussian noise with sigma = rel_sigma * max|ksp|."""
    rng = np.random.default_rng(seed + 104729)
    s = rel_sigma * float(np.max(np.abs(ksp)))
    n = rng.standard_normal(ksp.shape) + 1j * rng.standard_normal(ksp.shape)
    return (ksp + s / np.sqrt(2.0) * n).astype(np.complex64)


# --------------------------------------------------------------------------
# The benchmark workloads (sizes, seeds and trajectories in one place, for
# the benchmark and the full-size parity tests)
# --------------------------------------------------------------------------

def config_radial2d():
    """2-D golden-angle radial: 256^2 image of 10 ellipses, 8 Gaussian ring
    coils, 201 spokes x 512 samples (2x readout oversampling), lambda 0.01,
    30 CG iterations."""
    n, C, spokes, R = 256, 8, 201, 512
    return dict(name="2d_golden_radial_256^2_8coils_201x512", img=(n, n, 1), C=C, R=R, S=spokes,
                traj=radial_traj(R, spokes), x=ellipse_phantom(n, nell=10, seed=1),
                maps=gaussian_coils((1, n, n), C, seed=1), lam=0.01, iters=30, groups=2, seed=1)


def config_kooshball():
    """3-D radial kooshball: 128^3 image of 16 ellipsoids, 16 coils with 3-D
    Gaussian maps, 8192 phyllotaxis spokes x 256 samples, lambda 0.01,
    20 CG iterations."""
    n, C, spokes, R = 128, 16, 8192, 256
    return dict(name="3d_radial_kooshball_128^3_16coils_8192x256", img=(n, n, n), C=C, R=R, S=spokes,
                traj=phyllotaxis_kooshball(spokes=spokes, readout=R), x=ellipsoid_phantom(n, nell=16, seed=2),
                maps=gaussian_coils((n, n, n), C, seed=2), lam=0.01, iters=20, groups=16, seed=2)


def config_spiral():
    """2-D variable-density spiral: 320^2 image of 10 ellipses, 32 coils,
    24 interleaves x 4096 samples with density ~ 1 / (1 + 8 |k|), lambda
    0.01, 30 CG iterations."""
    n, C = 320, 32
    return dict(name="2d_vd_spiral_320^2_32coils_24x4096", img=(n, n, 1), C=C, R=4096, S=24,
                traj=vd_spiral(interleaves=24, samples=4096, n=n), x=ellipse_phantom(n, nell=10, seed=3),
                maps=gaussian_coils((1, n, n), C, seed=3), lam=0.01, iters=30, groups=16, seed=3)


def batched_slices(nslices=96, seed=4):
    """Uneven 2-D CG-SENSE problems: matrix n in U{128,192,256,320,384}, coils
    in U{8,16,32}, golden-angle radial spokes in U[21,201] with readout 2n;
    per slice an ellipse phantom and Gaussian coil maps of its own seed."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(nslices):
        n = int(rng.choice([128, 192, 256, 320, 384]))
        C = int(rng.choice([8, 16, 32]))
        spokes = int(rng.integers(21, 202))
        out.append(dict(n=n, C=C, R=2 * n, S=spokes, seed=1000 + s))
    return out
