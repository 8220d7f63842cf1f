"""CPU oracle for the NUFFT + CG-SENSE path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package; the product path in ``bart_amd`` never does.

Three oracles live here:

* ``liboracle.so`` (``bart_oracle.c``): a C/OpenMP restatement of BART's CPU
  NUFFT (noncart/nufft.c, noncart/grid.c), Toeplitz PSF and CG
  (iter/italgos.c:497-580).  Wrapped by :class:`Nufft` below.
* :func:`nudft_forward` / :func:`nudft_adjoint`: FP64 direct NUDFT with BART's
  sign and centring conventions (noncart/nudft.c:17-190).
* ``oracle/cnn.py``: FP64 statement of the warm-start network's layers and
  of the whole network's backpropagation (no counterpart in BART).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KB_SIZE = 1000


def build():
    """Compile liboracle.so from bart_oracle.c (gcc + OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        c_long3 = ctypes.c_long * 3
        vp = ctypes.c_void_p
        L.orc_bessel_i0.restype = ctypes.c_double
        L.orc_bessel_i0.argtypes = [ctypes.c_double]
        L.orc_calc_beta.restype = ctypes.c_double
        L.orc_calc_beta.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_kb_table.argtypes = [ctypes.c_double, vp]
        L.orc_rolloff.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_double, c_long3, vp]
        L.orc_fft.argtypes = [c_long3, ctypes.c_long, ctypes.c_int, vp]
        L.orc_fftmod_table.argtypes = [c_long3, ctypes.c_bool, vp]
        L.orc_linphase_table.argtypes = [c_long3, ctypes.c_int, vp, vp]
        L.orc_nufft_create.restype = vp
        L.orc_nufft_create.argtypes = [c_long3, ctypes.c_long, vp, vp, ctypes.c_int,
                                       ctypes.c_bool, ctypes.c_bool, ctypes.c_float, ctypes.c_float]
        L.orc_nufft_free.argtypes = [vp]
        L.orc_nufft_forward.argtypes = [vp, vp, vp]
        L.orc_nufft_adjoint.argtypes = [vp, vp, vp]
        L.orc_nufft_normal.argtypes = [vp, vp, vp]
        L.orc_nufft_get_psf.argtypes = [vp, vp]
        L.orc_nufft_set_psf.argtypes = [vp, vp]
        L.orc_nufft_get_linphase.argtypes = [vp, vp]
        L.orc_nufft_beta.restype = ctypes.c_double
        L.orc_nufft_beta.argtypes = [vp]
        L.orc_nufft_subgrids.restype = ctypes.c_int
        L.orc_nufft_subgrids.argtypes = [vp]
        L.orc_nufft_compute_psf.argtypes = [c_long3, ctypes.c_long, vp, vp, vp]
        L.orc_nufft_compute_psf_vals.argtypes = [c_long3, ctypes.c_long, vp, vp, vp, vp]
        L.orc_nufft_normal_basis.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.orc_chain_create.restype = vp
        L.orc_chain_create.argtypes = [c_long3, ctypes.c_long, vp, vp, ctypes.c_int, ctypes.c_bool, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_bool]
        L.orc_chain_free.argtypes = [vp]
        L.orc_chain_forward.argtypes = [vp, vp, vp]
        L.orc_chain_adjoint.argtypes = [vp, vp, vp]
        L.orc_psf_grid_fp64.argtypes = [c_long3, ctypes.c_long, vp, vp, ctypes.c_int, vp]
        L.orc_nufft_grid_fp64.argtypes = [vp, ctypes.c_int, vp, vp]
        L.orc_sense_forward.argtypes = [vp, vp, vp, vp]
        L.orc_sense_adjoint.argtypes = [vp, vp, vp, vp]
        L.orc_sense_normal.argtypes = [vp, vp, vp, vp]
        L.orc_cgsense.restype = ctypes.c_int
        L.orc_cgsense.argtypes = [vp, vp, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                                  vp, vp, vp, vp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _dims3(dims):
    d = list(dims) + [1] * (3 - len(dims))
    return (ctypes.c_long * 3)(*d)


def bessel_i0(x):
    return lib().orc_bessel_i0(float(x))


def calc_beta(os_=2.0, width=6.0):
    return lib().orc_calc_beta(os_, width)


def kb_table(beta):
    t = np.zeros(KB_SIZE + 1, np.float32)
    lib().orc_kb_table(beta, _ptr(t))
    return t


def rolloff(dims, os_=2.0, width=6.0, beta=None):
    if beta is None:
        beta = calc_beta(os_, width)
    d = list(dims) + [1] * (3 - len(dims))
    out = np.zeros(d[::-1], np.complex64)
    lib().orc_rolloff(os_, width, beta, _dims3(d), _ptr(out))
    return out


def fft(x, sign=-1):
    """Unnormalised FFT over the last (up to 3) axes of a C-ordered array whose
    trailing axes are (Z, Y, X); leading axes are batch."""
    x = _c64(x).copy()
    shp = x.shape
    dims = list(shp[::-1][:3]) + [1] * max(0, 3 - len(shp))
    batch = int(np.prod(shp[:-3])) if len(shp) > 3 else 1
    lib().orc_fft(_dims3(dims), batch, sign, _ptr(x))
    return x


def fftmod_table(dims, inv=False):
    d = list(dims) + [1] * (3 - len(dims))
    out = np.zeros(d[::-1], np.complex64)
    lib().orc_fftmod_table(_dims3(d), inv, _ptr(out))
    return out


def linphase_table(dims, s, roll=None):
    d = list(dims) + [1] * (3 - len(dims))
    out = np.zeros(d[::-1], np.complex64)
    r = _c64(roll) if roll is not None else None
    lib().orc_linphase_table(_dims3(d), s, _ptr(r), _ptr(out))
    return out


def compute_psf(dims, traj, weights=None):
    """PSF [S, Z, Y, X] as BART's compute_psf2 (noncart/nufft.c:727-776).
    traj: complex64 [M, 3] (real part = coordinate)."""
    d = list(dims) + [1] * (3 - len(dims))
    S = 2 ** sum(1 for v in d if v > 1)
    tr = _c64(traj).reshape(-1, 3)
    w = _c64(weights).reshape(-1) if weights is not None else None
    out = np.zeros([S] + d[::-1], np.complex64)
    lib().orc_nufft_compute_psf(_dims3(d), tr.shape[0], _ptr(tr), _ptr(w), _ptr(out))
    return out


def compute_psf_fp64(dims, traj, weights=None):
    """The PSF of :func:`compute_psf` (noncart/nufft.c:727-776 with
    :535-586) evaluated in complex128: gridding with double products and
    sums (``orc_psf_grid_fp64``), then conj(fftmod), unnormalised inverse
    FFT, conj(linphase_s) summed over the sub-grids, rolloff, and the
    centred unitary FFT and decomposition, all in complex128 on the FP32
    tables taken as exact.  It differs from the FP32 PSF only by the FP32
    rounding, which grows with the number of samples gridded onto one
    point (the dense centre of radial trajectories).  Returns [S, Z, Y, X]
    complex128."""
    import scipy.fft as sf
    d = list(dims) + [1] * (3 - len(dims))
    d2 = [1 if v == 1 else 2 * v for v in d]
    tr = _c64(traj).reshape(-1, 3)
    w = _c64(weights).reshape(-1) if weights is not None else None
    M = tr.shape[0]
    S2 = 2 ** sum(1 for v in d2 if v > 1)
    N2 = int(np.prod(d2))
    axes = tuple(a for a in range(3) if d2[::-1][a] > 1)
    beta = float(np.float32(calc_beta(2.0, 6.0)))
    fm = fftmod_table(d2).astype(np.complex128)
    acc = np.zeros(d2[::-1], np.complex128)
    g = np.zeros(d2[::-1], np.complex128)
    for s in range(S2):
        lib().orc_psf_grid_fp64(_dims3(d), M, _ptr(tr), _ptr(w), s, _ptr(g))
        v = sf.ifftn(g * np.conj(fm), axes=axes, workers=-1) * N2
        v *= np.conj(linphase_table(d2, s).astype(np.complex128))
        acc += v
    del g
    acc *= rolloff(d2, 2.0, 6.0, beta).astype(np.complex128)
    acc *= fm
    acc = sf.fftn(acc, axes=axes, overwrite_x=True, workers=-1)
    acc *= fm / np.sqrt(N2)
    # md_decompose: psf[s][j] = psft[2 j + r(s)] (r(s) bit i = shift of dim i)
    S = 2 ** sum(1 for v in d if v > 1)
    out = np.zeros((S,) + tuple(d[::-1]), np.complex128)
    for s in range(S):
        r, idx = [], s
        for v in d:
            f = 2 if v > 1 else 1
            r.append(idx % f)
            idx //= f
        sl = tuple(slice(r[i], None, 2) if d[i] > 1 else slice(None) for i in (2, 1, 0))
        out[s] = acc[sl]
    return out


class Nufft:
    """BART-style NUFFT on the CPU.

    dims: image dims (X, Y, Z); traj: complex64 [M, 3]; coil images are
    arrays [C, Z, Y, X]; k-space arrays are [C, M].
    """

    def __init__(self, dims, traj, ncoils, weights=None, toeplitz=True, periodic=False,
                 width=6.0, os_=2.0):
        self.dims = tuple(list(dims) + [1] * (3 - len(dims)))
        self.traj = _c64(traj).reshape(-1, 3)
        self.M = self.traj.shape[0]
        self.C = int(ncoils)
        self.weights = _c64(weights).reshape(-1) if weights is not None else None
        self.toeplitz = toeplitz
        self.h = lib().orc_nufft_create(_dims3(self.dims), self.M, _ptr(self.traj),
                                        _ptr(self.weights), self.C, toeplitz, periodic,
                                        width, os_)
        if not self.h:
            raise ValueError("oracle supports os = 2 only")
        self.S = lib().orc_nufft_subgrids(self.h)
        self.N = int(np.prod(self.dims))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_nufft_free(self.h)
            self.h = None

    @property
    def cim_shape(self):
        return (self.C,) + self.dims[::-1]

    def forward(self, cim):
        cim = _c64(cim).reshape(self.cim_shape)
        out = np.zeros((self.C, self.M), np.complex64)
        lib().orc_nufft_forward(self.h, _ptr(out), _ptr(cim))
        return out

    def adjoint(self, ksp):
        ksp = _c64(ksp).reshape(self.C, self.M)
        out = np.zeros(self.cim_shape, np.complex64)
        lib().orc_nufft_adjoint(self.h, _ptr(out), _ptr(ksp))
        return out

    def normal(self, cim):
        assert self.toeplitz
        cim = _c64(cim).reshape(self.cim_shape)
        out = np.zeros(self.cim_shape, np.complex64)
        lib().orc_nufft_normal(self.h, _ptr(out), _ptr(cim))
        return out

    def psf(self):
        out = np.zeros((self.S,) + self.dims[::-1], np.complex64)
        lib().orc_nufft_get_psf(self.h, _ptr(out))
        return out

    def set_psf(self, psf):
        psf = _c64(psf)
        lib().orc_nufft_set_psf(self.h, _ptr(psf))

    def linphase(self):
        out = np.zeros((self.S,) + self.dims[::-1], np.complex64)
        lib().orc_nufft_get_linphase(self.h, _ptr(out))
        return out

    # SENSE model (maps: [C, Z, Y, X], image [Z, Y, X])
    def sense_forward(self, maps, img):
        maps = _c64(maps).reshape(self.cim_shape)
        img = _c64(img).reshape(self.dims[::-1])
        out = np.zeros((self.C, self.M), np.complex64)
        lib().orc_sense_forward(self.h, _ptr(maps), _ptr(out), _ptr(img))
        return out

    def sense_adjoint(self, maps, ksp):
        maps = _c64(maps).reshape(self.cim_shape)
        ksp = _c64(ksp).reshape(self.C, self.M)
        out = np.zeros(self.dims[::-1], np.complex64)
        lib().orc_sense_adjoint(self.h, _ptr(maps), _ptr(out), _ptr(ksp))
        return out

    def sense_normal(self, maps, img):
        maps = _c64(maps).reshape(self.cim_shape)
        img = _c64(img).reshape(self.dims[::-1])
        out = np.zeros(self.dims[::-1], np.complex64)
        lib().orc_sense_normal(self.h, _ptr(maps), _ptr(out), _ptr(img))
        return out

    @property
    def beta(self):
        """Kaiser-Bessel beta of the operator's table and rolloff."""
        return lib().orc_nufft_beta(self.h)

    def adjoint_fp64(self, ksp):
        """The adjoint NUFFT (noncart/nufft.c:1462-1540) in complex128:
        gridding with double products and sums (``orc_nufft_grid_fp64``),
        then conj(fftmod), the unnormalised inverse FFT and conj(linphase_s)
        summed over the sub-grids (times the rolloff of a Toeplitz
        operator, whose linear phases do not hold it), on the FP32 tables
        taken as exact.  It differs from :meth:`adjoint` only by the FP32
        rounding, which grows with the number of samples gridded onto one
        point (the dense centre of radial trajectories).  Returns
        [C, Z, Y, X] complex128."""
        import scipy.fft as sf
        ksp = _c64(ksp).reshape(self.C, self.M)
        lph = self.linphase()
        fm = fftmod_table(self.dims).astype(np.complex128)
        axes = tuple(a for a in (1, 2, 3) if self.cim_shape[a] > 1)
        acc = np.zeros(self.cim_shape, np.complex128)
        g = np.zeros(self.cim_shape, np.complex128)
        for s in range(self.S):
            lib().orc_nufft_grid_fp64(self.h, s, _ptr(ksp), _ptr(g))
            v = sf.ifftn(g * np.conj(fm)[None], axes=axes, workers=-1) * self.N
            v *= np.conj(lph[s].astype(np.complex128))[None]
            acc += v
        if self.toeplitz:
            acc *= rolloff(self.dims, 2.0, 6.0, self.beta).astype(np.complex128)[None]
        return acc

    def sense_adjoint_fp64(self, maps, ksp):
        """S^H of :meth:`adjoint_fp64`: sum over coils of conj(map) times the
        coil image, in complex128."""
        m = np.asarray(maps).astype(np.complex128).reshape(self.cim_shape)
        return np.sum(np.conj(m) * self.adjoint_fp64(ksp), axis=0)

    def normal_fp64(self, cim, psf=None):
        """The Toeplitz normal operator of the coil images (no maps) in
        complex128: per coil sum_s conj(L_s) IFFT(P_s FFT(L_s x))
        (noncart/nufft.c:1546-1625) on this operator's linear phases taken
        as exact, with this operator's PSF or the one given (for example
        :func:`compute_psf_fp64`)."""
        import scipy.fft as sf
        assert self.toeplitz
        lph = self.linphase().astype(np.complex128)
        P = self.psf().astype(np.complex128) if psf is None else np.asarray(psf, np.complex128)
        x = np.asarray(cim).astype(np.complex128).reshape(self.cim_shape)
        axes = tuple(a for a in (1, 2, 3) if self.cim_shape[a] > 1)
        acc = np.zeros(self.cim_shape, np.complex128)
        for s in range(self.S):
            g = sf.fftn(x * lph[s][None], axes=axes, workers=-1)
            g *= P[s][None]
            g = sf.ifftn(g, axes=axes, workers=-1, overwrite_x=True) * self.N
            acc += g * np.conj(lph[s])[None]
        return acc

    def sense_normal_fp64(self, maps, img, psf=None):
        """S^H T S in complex128: the Toeplitz normal operator
        (noncart/nufft.c:1546-1625, sum_s conj(L_s) IFFT(P_s FFT(L_s x)) per
        coil, then the coil sum of linops/fmac.c) evaluated in double
        precision on this operator's own tables (linear phases, PSF) and the
        maps, each taken as exact (or with the PSF given, for example
        :func:`compute_psf_fp64`).  The reference for the FP32 rounding of
        CG: an FP32 solver's deviation from CG on this operator is the
        rounding it accumulates."""
        import scipy.fft as sf
        assert self.toeplitz
        if not hasattr(self, "_lph64"):
            self._lph64 = self.linphase().astype(np.complex128)
            self._psf64 = self.psf().astype(np.complex128)
        P = self._psf64 if psf is None else np.asarray(psf, np.complex128)
        m = np.asarray(maps).astype(np.complex128).reshape(self.cim_shape)
        x = np.asarray(img).astype(np.complex128).reshape(self.dims[::-1])
        axes = tuple(a for a in (1, 2, 3) if self.cim_shape[a] > 1)
        cim = m * x[None]
        acc = np.zeros(self.cim_shape, np.complex128)
        for s in range(self.S):
            g = sf.fftn(cim * self._lph64[s][None], axes=axes, workers=-1)
            g *= P[s][None]
            g = sf.ifftn(g, axes=axes, workers=-1) * self.N
            acc += g * np.conj(self._lph64[s])[None]
        return np.sum(np.conj(m) * acc, axis=0)

    def cgsense(self, maps, b, lam, maxiter, tol=0.0, x0=None):
        """CG on (A^H A + lam I) x = b as BART's conjgrad.  Returns
        (x, iterations, residual trace, final residual norm)."""
        maps = _c64(maps).reshape(self.cim_shape)
        b = _c64(b).reshape(self.dims[::-1])
        x = np.zeros(self.dims[::-1], np.complex64) if x0 is None else _c64(x0).reshape(self.dims[::-1]).copy()
        trace = np.zeros(max(1, maxiter), np.float32)
        fin = np.zeros(1, np.float32)
        it = lib().orc_cgsense(self.h, _ptr(maps), lam, maxiter, tol, _ptr(x), _ptr(b),
                               _ptr(trace), _ptr(fin))
        return x, it, trace[:it], float(fin[0])


class BasisNufft:
    """BART's NUFFT with a subspace basis on the CPU
    (noncart/nufft.c:470-505, 1018-1036, 1450-1453, 1483-1487, 1546-1625).

    traj: complex64 [T or 1, Mf, 3] (one trajectory per frame, or one for
    all); basis: [T, K] (B[t][k], BART dims [1 1 1 1 1 T K]); weights:
    [T or 1, Mf] or None.  Coefficient coil images are [K, C, Z, Y, X];
    frame k-space is [T, C, Mf].  The coefficient NUFFT is the plain one
    with C K channels (c + C k) over the T Mf samples of all frames.
    """

    def __init__(self, dims, traj, basis, ncoils, weights=None, toeplitz=True, periodic=False):
        self.B = _c64(basis)
        self.T, self.K = self.B.shape
        tr = _c64(traj)
        tr = tr.reshape((1, -1, 3)) if tr.ndim == 2 else tr
        self.Mf = tr.shape[1]
        tr = np.broadcast_to(tr, (self.T, self.Mf, 3)).reshape(-1, 3)
        self.traj = np.ascontiguousarray(tr)
        w = None
        if weights is not None:
            w = _c64(weights).reshape(-1, self.Mf)
            w = np.ascontiguousarray(np.broadcast_to(w, (self.T, self.Mf)).reshape(-1))
        self.weights = w
        self.C = int(ncoils)
        self.dims = tuple(list(dims) + [1] * (3 - len(dims)))
        # in Toeplitz mode the linear phases exclude the rolloff, which the
        # forward and adjoint apply separately (as BART's plan does); the
        # plain PSF the Toeplitz plan also computes is not used
        self.nf = Nufft(self.dims, self.traj, self.C * self.K, weights=w, toeplitz=toeplitz, periodic=periodic)
        self.S = self.nf.S
        self.N = self.nf.N
        self.psfB = self.compute_psf() if toeplitz else None

    @property
    def cim_shape(self):
        return (self.K, self.C) + self.dims[::-1]

    def forward(self, cim):
        """[K, C, ...] -> [T, C, Mf]: dst[t][c] = sum_k B[t][k] NUFFT(x_kc)[t]
        (md_ztenmul, nufft.c:1450-1453; weights inside the plain NUFFT)."""
        tmp = self.nf.forward(_c64(cim).reshape(self.K * self.C, -1)).reshape(self.K, self.C, self.T, self.Mf)
        out = np.zeros((self.T, self.C, self.Mf), np.complex64)
        for k in range(self.K):
            out += self.B[:, k][:, None, None] * tmp[k].transpose(1, 0, 2)
        return out

    def adjoint(self, ksp):
        """[T, C, Mf] -> [K, C, ...]: x conj(B[t][k]) (md_ztenmulc,
        nufft.c:1483-1487), then the plain adjoint NUFFT."""
        ksp = _c64(ksp).reshape(self.T, self.C, self.Mf)
        tmp = np.zeros((self.K, self.C, self.T, self.Mf), np.complex64)
        for k in range(self.K):
            tmp[k] = (ksp * np.conj(self.B[:, k])[:, None, None]).transpose(1, 0, 2)
        return self.nf.adjoint(tmp.reshape(self.K * self.C, -1)).reshape(self.cim_shape)

    def compute_psf(self):
        """[K K, S, Z, Y, X]: entry k K + k' = the PSF of |w|^2 conj(B[t][k])
        B[t][k'] (compute_psf_int with compute_square_basis,
        nufft.c:470-505, 535-586)."""
        d = list(self.dims)
        out = np.zeros((self.K * self.K, self.S) + tuple(d[::-1]), np.complex64)
        t_of = np.repeat(np.arange(self.T), self.Mf)
        for k in range(self.K):
            for q in range(self.K):
                vals = np.ascontiguousarray((np.conj(self.B[t_of, k]) * self.B[t_of, q]).astype(np.complex64))
                o = np.zeros((self.S,) + tuple(d[::-1]), np.complex64)
                lib().orc_nufft_compute_psf_vals(_dims3(d), self.traj.shape[0], _ptr(self.traj),
                                                 _ptr(self.weights), _ptr(vals), _ptr(o))
                out[k * self.K + q] = o
        return out

    def normal(self, cim):
        """Toeplitz normal operator with the K x K PSF (nufft.c:1546-1625)."""
        cim = _c64(cim).reshape(self.cim_shape)
        out = np.zeros(self.cim_shape, np.complex64)
        psf = np.ascontiguousarray(self.psfB)
        lib().orc_nufft_normal_basis(self.nf.h, self.K, _ptr(psf), _ptr(out), _ptr(cim))
        return out

    def sense_forward(self, maps, img):
        """img [K, ...], maps [C, ...] -> [T, C, Mf] (maps2, sense/modelnc.c:57-59)."""
        img = _c64(img).reshape((self.K, 1) + self.dims[::-1])
        return self.forward(img * _c64(maps).reshape((1, self.C) + self.dims[::-1]))

    def sense_adjoint(self, maps, ksp):
        cim = self.adjoint(ksp)
        return np.sum(cim * np.conj(_c64(maps).reshape((1, self.C) + self.dims[::-1])), axis=1)

    def sense_normal(self, maps, img):
        m = _c64(maps).reshape((1, self.C) + self.dims[::-1])
        img = _c64(img).reshape((self.K, 1) + self.dims[::-1])
        cim = self.normal(img * m)
        return np.sum(cim * np.conj(m), axis=1)


class ChainNufft:
    """BART's chained NUFFT for oversampling factors other than 2
    (noncart/nufft_chain.c:192-244): rolloff into the centre of the
    round(os n) grid, centred unitary FFT, KB interpolation with beta =
    calc_beta(os, width), weights times os^(d/2).  Arrays as :class:`Nufft`.
    pad=False: the non-decomposed NUFFT (decomp = false, nufft -1,
    noncart/nufft.c:898-935): the image is the oversampled grid itself, the
    rolloff is taken with os = 1 over it."""

    def __init__(self, dims, traj, ncoils, weights=None, periodic=False, width=6.0, os_=1.5, pad=True):
        self.dims = tuple(list(dims) + [1] * (3 - len(dims)))
        self.traj = _c64(traj).reshape(-1, 3)
        self.M = self.traj.shape[0]
        self.C = int(ncoils)
        self.weights = _c64(weights).reshape(-1) if weights is not None else None
        self.h = lib().orc_chain_create(_dims3(self.dims), self.M, _ptr(self.traj), _ptr(self.weights), self.C,
                                        periodic, width, os_, pad)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_chain_free(self.h)
            self.h = None

    @property
    def cim_shape(self):
        return (self.C,) + self.dims[::-1]

    def forward(self, cim):
        cim = _c64(cim).reshape(self.cim_shape)
        out = np.zeros((self.C, self.M), np.complex64)
        lib().orc_chain_forward(self.h, _ptr(out), _ptr(cim))
        return out

    def adjoint(self, ksp):
        ksp = _c64(ksp).reshape(self.C, self.M)
        out = np.zeros(self.cim_shape, np.complex64)
        lib().orc_chain_adjoint(self.h, _ptr(out), _ptr(ksp))
        return out


def precond_strang(nf):
    """Strang's circulant preconditioner of the Toeplitz operator nf
    (noncart/precond.c:46-71), complex128 on nf's FP32 PSF and linear phases:
    pre = |fftuc(sum_s ifftuc(psf_s) conj(linphase_s))| + 1e-3 over the
    image dims.  Returns (pre [Z, Y, X] real, apply) with apply(x) =
    IFFT(FFT(x) / pre) per coil, unnormalised FFTs (precond.c:92-106)."""
    assert nf.toeplitz
    d = nf.dims
    axes = tuple(a for a in range(-3, 0) if d[::-1][a] > 1)
    fm = fftmod_table(d).astype(np.complex128)
    N = float(np.prod(d))
    psf = nf.psf().astype(np.complex128)
    lph = nf.linphase().astype(np.complex128)
    acc = np.zeros(d[::-1], np.complex128)
    for s in range(nf.S):
        t = np.fft.ifftn(np.conj(fm) * psf[s], axes=axes) * N     # unnormalised inverse FFT
        acc += (np.conj(fm) * t / np.sqrt(N)) * np.conj(lph[s])
    pre = np.abs(fm * np.fft.fftn(fm * acc, axes=axes) / np.sqrt(N)) + 1e-3

    def apply(x):
        x = np.asarray(x, np.complex128)
        ax = tuple(range(x.ndim - 3, x.ndim))
        ax = tuple(a for a in ax if x.shape[a] > 1)
        return np.fft.ifftn(np.fft.fftn(x, axes=ax) / pre, axes=ax) * N

    return pre, apply


def cg_fp64(normal, b, lam, maxiter, tol=0.0, x0=None, snapshots=None):
    """CG on (A^H A + lam I) x = b in complex128, step by step as BART's
    conjgrad (iter/italgos.c:497-580): the inner products of the float
    vectors are Re<u, v> (num/vecops.c:89-124), stop when
    ||r||^2 <= (tol ||b||)^2 (iter/iter2.c:227-270) or p^H A p = 0.
    normal: function image -> A^H A image (complex128).  Returns (x, trace
    of ||r|| after each iteration); with snapshots (a dict whose keys are
    iteration counts) the iterate after each of those iterations is stored
    under its key (0: the start)."""
    b = np.asarray(b, np.complex128)
    shape = b.shape
    b = b.ravel()
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, np.complex128).ravel().copy()
    eps2 = (tol * np.linalg.norm(b)) ** 2
    r = b - (normal(x.reshape(shape)).ravel() + lam * x)
    p = r.copy()
    rsold = float(np.vdot(r, r).real)
    trace = []
    if snapshots is not None and 0 in snapshots:
        snapshots[0] = x.reshape(shape).copy()
    if rsold == 0.0:
        return x.reshape(shape), np.array(trace)
    for _ in range(maxiter):
        Ap = normal(p.reshape(shape)).ravel() + lam * p
        pAp = float(np.vdot(p, Ap).real)
        if pAp == 0.0:
            break
        alpha = rsold / pAp
        x += alpha * p
        r -= alpha * Ap
        rsnew = float(np.vdot(r, r).real)
        trace.append(np.sqrt(rsnew))
        if snapshots is not None and len(trace) in snapshots:
            snapshots[len(trace)] = x.reshape(shape).copy()
        if rsnew <= eps2:
            break
        p = r + (rsnew / rsold) * p
        rsold = rsnew
    return x.reshape(shape), np.array(trace)


# --------------------------------------------------------------------------
# FP64 direct NUDFT (noncart/nudft.c:17-190): coordinates are negated in
# get_coord, the phase is linear_phase centred at dims/2 (num/filter.c:241-265)
# and the result is scaled by 1/sqrt(N).
# --------------------------------------------------------------------------

def _phase_matrix(dims, traj, sign):
    d = list(dims) + [1] * (3 - len(dims))
    k = np.real(np.asarray(traj, np.complex128)).reshape(-1, 3)
    ph = np.zeros((k.shape[0],) + tuple(d[::-1]), np.float64)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in d[::-1]], indexing="ij")
    for ax in range(3):
        n = d[ax]
        if n == 1:
            continue
        pos = grids[2 - ax] - n / 2.0
        ph += (k[:, ax] / n).reshape((-1, 1, 1, 1)) * pos[None]
    return np.exp(sign * 2j * np.pi * ph)


def nudft_forward(dims, traj, img, chunk=2048):
    """img [..., Z, Y, X] -> ksp [..., M] in float64/complex128."""
    d = list(dims) + [1] * (3 - len(dims))
    img = np.asarray(img, np.complex128).reshape((-1,) + tuple(d[::-1]))
    tr = np.asarray(traj).reshape(-1, 3)
    out = np.zeros((img.shape[0], tr.shape[0]), np.complex128)
    N = int(np.prod(d))
    for i in range(0, tr.shape[0], chunk):
        E = _phase_matrix(d, tr[i:i + chunk], -1.0).reshape(-1, N)
        out[:, i:i + chunk] = img.reshape(img.shape[0], N) @ E.T
    return out / np.sqrt(N)


def nudft_adjoint(dims, traj, ksp, chunk=2048):
    """ksp [..., M] -> img [..., Z, Y, X] in complex128."""
    d = list(dims) + [1] * (3 - len(dims))
    tr = np.asarray(traj).reshape(-1, 3)
    ksp = np.asarray(ksp, np.complex128).reshape(-1, tr.shape[0])
    N = int(np.prod(d))
    out = np.zeros((ksp.shape[0], N), np.complex128)
    for i in range(0, tr.shape[0], chunk):
        E = _phase_matrix(d, tr[i:i + chunk], +1.0).reshape(-1, N)
        out += ksp[:, i:i + chunk] @ E
    return (out / np.sqrt(N)).reshape((ksp.shape[0],) + tuple(d[::-1]))


def nrmse(ref, x):
    """BART's nrmse (src/nrmse.c, md_znrmse): ||ref - x|| / ||ref||."""
    ref = np.asarray(ref, np.complex128)
    x = np.asarray(x, np.complex128)
    return float(np.linalg.norm((ref - x).ravel()) / np.linalg.norm(ref.ravel()))
