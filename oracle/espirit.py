"""CPU restatement of BART's ESPIRiT coil-sensitivity calibration
(src/ecalib.c, calib/calib.c, calib/calmat.c, linops/casorati.c,
misc/mri2.c, misc/resize.c, misc/utils.c, calib/cc.c, num/linalg.c) in
float64 numpy.  Test infrastructure: the library's GPU ecalib is checked
against it.

Array layout: numpy C order with BART's dims reversed, k-space
[C, Z, Y, X] for BART [X Y Z C]; maps [M, C, Z, Y, X] for BART
[X Y Z C M].

Parity notes: the eigenvector of the channel covariance that fixes the
phase of the maps (calib2, calib/calib.c:737-746 with scc, calib/cc.c:45-63)
is only defined up to a global phase, which BART leaves to LAPACK; here
its first entry is made real and non-negative, so maps may differ from
BART's by one global phase per call (the reference's own phase tests use
scaled comparisons, tests/ecalib.mk).
"""
import numpy as np


def _fftmod_1d(n, inv=False):
    """num/fft.c:64-110: exp(+-2 pi i (j - n/2/2) (n/2)/n), exact quarters."""
    c1 = n // 2
    shift = c1 / n
    j = np.arange(n, dtype=np.float64)
    ph = (j - c1 / 2.0) * shift
    rem = ph - np.floor(ph)
    v = np.exp(2j * np.pi * rem * (-1.0 if inv else 1.0))
    v[rem == 0.0] = 1.0
    v[rem == 0.5] = -1.0
    v[rem == 0.25] = -1j if inv else 1j
    v[rem == 0.75] = 1j if inv else -1j
    return v


def _mod_axis(x, axis, inv):
    n = x.shape[axis]
    if n == 1:
        return x
    shape = [1] * x.ndim
    shape[axis] = n
    return x * _fftmod_1d(n, inv).reshape(shape)


def _resize_center(x, shape):
    """md_resize_center: crop or zero-pad about the centre n/2."""
    out = np.zeros(shape, np.complex128)
    src, dst = [], []
    for n_in, n_out in zip(x.shape, shape):
        if n_out >= n_in:
            o = n_out // 2 - n_in // 2
            src.append(slice(0, n_in))
            dst.append(slice(o, o + n_in))
        else:
            o = n_in // 2 - n_out // 2
            src.append(slice(o, o + n_out))
            dst.append(slice(0, n_out))
    out[tuple(dst)] = x[tuple(src)]
    return out


def calib_geom(ksp, calsize):
    """misc/mri2.c:101-168 on the centre of calsize (y, z) x full readout
    (extract_calib2, :171-217): grow a centred box while every sample in it
    is non-zero in some coil, then move it along the readout to the
    position of maximal energy.  Returns the calibration data [C, z, y, x]."""
    C, Z, Y, X = ksp.shape
    dims = [X, Y, Z]
    tmpd = [X, min(calsize[1], Y), min(calsize[2], Z)]
    tpos = [(dims[i] - tmpd[i]) // 2 for i in range(3)]
    tmp = ksp[:, tpos[2]:tpos[2] + tmpd[2], tpos[1]:tpos[1] + tmpd[1], tpos[0]:tpos[0] + tmpd[0]]
    pat = (np.abs(tmp).sum(axis=0) != 0).astype(np.float64)   # estimate_pattern over coils
    cal = [1, 1, 1]
    pos = [(tmpd[i] - 1) // 2 for i in range(3)]
    stop = [False] * 3
    while not (stop[0] and stop[1] and stop[2]):
        for i in range(3):
            if cal[i] == tmpd[i] or cal[i] >= calsize[i]:
                stop[i] = True
            if stop[i]:
                continue
            cal[i] += 1
            pos[i] = (tmpd[i] - cal[i]) // 2
            blk = pat[pos[2]:pos[2] + cal[2], pos[1]:pos[1] + cal[1], pos[0]:pos[0] + cal[0]]
            si = np.float32(np.sqrt(np.float32(cal[0]) * np.float32(cal[1]) * np.float32(cal[2])))
            if si != np.float32(np.sqrt(np.sum(blk ** 2))):
                cal[i] -= 1
                pos[i] = (tmpd[i] - cal[i]) // 2
                stop[i] = True
    # calib_readout_pos: maximal energy along the readout
    best, beng = 0, 0.0
    for r in range(tmpd[0] - cal[0] + 1):
        e = np.sum(np.abs(tmp[:, pos[2]:pos[2] + cal[2], pos[1]:pos[1] + cal[1], r:r + cal[0]]) ** 2)
        if e > beng:
            best, beng = r, e
    pos[0] = best
    return np.ascontiguousarray(tmp[:, pos[2]:pos[2] + cal[2], pos[1]:pos[1] + cal[1], pos[0]:pos[0] + cal[0]])


def casorati_gram(cal, kdims):
    """covariance_function_fft (calib/calmat.c:181-186) = the Gram matrix
    A^H A of the calibration (Casorati) matrix: rows the block positions,
    columns the kernel entries (kx fastest, then ky, kz, coil)."""
    C, cz, cy, cx = cal.shape
    kx, ky, kz = kdims
    P = (cx - kx + 1) * (cy - ky + 1) * (cz - kz + 1)
    N = kx * ky * kz * C
    A = np.zeros((P, N), np.complex128)
    col = 0
    for c in range(C):
        for oz in range(kz):
            for oy in range(ky):
                for ox in range(kx):
                    A[:, col] = cal[c, oz:oz + cz - kz + 1, oy:oy + cy - ky + 1, ox:ox + cx - kx + 1].reshape(-1)
                    col += 1
    return A.conj().T @ A


def compute_kernels(G, threshold):
    """compute_kernels (calib/calib.c:882-1004): eigenvectors of the Gram
    matrix in descending order, singular values sqrt(eig), and the number
    kept: val_i / val_0 > sqrt(threshold) (number_of_kernels, :835-866).
    The kernels are the conjugated eigenvectors: lapack_eig reads BART's
    row-major matrix as its transpose, the conjugate of the Hermitian Gram
    matrix, and returns the eigenvectors in the rows BART then uses."""
    w, V = np.linalg.eigh(G)
    w = w[::-1]
    V = V[:, ::-1]
    val = np.sqrt(np.maximum(w, 0.0))
    n = int(np.sum(val / val[0] > np.sqrt(np.float32(threshold))))
    return V[:, :n].T.conj().copy(), val, n


def compute_imgcov(kerns, kdims, C):
    """compute_imgcov (calib/calib.c:1007-1094): kernels zero-padded to
    twice their size, ifftmod and unnormalised inverse FFT, per point the
    C x C Gram matrix sum_k conj(K[c1, k]) K[c2, k] / (kx ky kz xh yh zh),
    returned full [Z, Y, X, C, C] (upper triangle the conjugate)."""
    kx, ky, kz = kdims
    xh, yh, zh = [(1 if k == 1 else 2 * k) for k in (kx, ky, kz)]
    nk = kerns.shape[0]
    K = kerns.reshape(nk, C, kz, ky, kx)
    Kp = _resize_center(K, (nk, C, zh, yh, xh))
    for ax in (2, 3, 4):
        Kp = _mod_axis(Kp, ax, inv=True)
    Kp = np.fft.ifftn(Kp, axes=(2, 3, 4)) * (xh * yh * zh)
    # [nk, C, z, y, x] -> per point data[c][k]
    D = np.transpose(Kp, (2, 3, 4, 1, 0))
    Gr = np.einsum("zyxik,zyxjk->zyxij", D.conj(), D)
    return Gr / float(kx * ky * kz * xh * yh * zh)


def sinc_zeropad(x, out_shape):
    """misc/resize.c:234-292 over the spatial axes of x [Z, Y, X, ...]:
    per resized axis (z first), fftmod, FFT, fftmod, centred zero-padding,
    ifftmod, unnormalised inverse FFT, ifftmod."""
    for ax in (0, 1, 2):
        n_in, n_out = x.shape[ax], out_shape[ax]
        if n_in == n_out:
            continue
        t = _mod_axis(x, ax, False)
        t = np.fft.fft(t, axis=ax)
        t = _mod_axis(t, ax, False)
        shape = list(t.shape)
        shape[ax] = n_out
        t = _resize_center(t, tuple(shape))
        t = _mod_axis(t, ax, True)
        t = np.fft.ifft(t, axis=ax) * n_out
        x = _mod_axis(t, ax, True)
    return x


def orthiter(M, H, iters):
    """orthiter (num/linalg.c:1296-1314): rows start as the identity; each
    step multiplies them by H from the right and re-orthonormalises with
    gram_schmidt (:1235-1251), which keeps the last row free and removes
    from every row the components of the rows after it.  Returns (rows,
    vals)."""
    N = H.shape[0]
    out = np.eye(M, N, dtype=np.complex128)
    vals = np.zeros(M)
    for _ in range(iters):
        out = out @ H
        for r in range(M - 1, -1, -1):
            for j in range(r + 1, M):
                s = np.sum(out[r] * np.conj(out[j]))
                out[r] = out[r] - s * out[j]
            vals[r] = np.sqrt(np.real(np.sum(out[r] * np.conj(out[r]))))
            out[r] = out[r] / vals[r]
    return out, vals


def orthiter_batch(M, H, iters):
    """:func:`orthiter` for a stack of matrices H [P, N, N]."""
    P, N, _ = H.shape
    out = np.broadcast_to(np.eye(M, N, dtype=np.complex128), (P, M, N)).copy()
    vals = np.zeros((P, M))
    for _ in range(iters):
        out = out @ H
        for r in range(M - 1, -1, -1):
            for j in range(r + 1, M):
                sc = np.sum(out[:, r] * np.conj(out[:, j]), axis=1)
                out[:, r] = out[:, r] - sc[:, None] * out[:, j]
            vals[:, r] = np.sqrt(np.real(np.sum(out[:, r] * np.conj(out[:, r]), axis=1)))
            out[:, r] = out[:, r] / vals[:, r][:, None]
    return out, vals


def scc_rot(cal):
    """The principal component of the channel covariance of the calibration
    data (scc, calib/cc.c:45-63, as used by calib2 :737-746): first entry
    made real and non-negative (see the module notes)."""
    C = cal.shape[0]
    D = cal.reshape(C, -1)
    T = D.conj() @ D.T          # gram_matrix: T[i][j] = sum conj(d_i) d_j
    w, V = np.linalg.eigh(np.conj(T))
    v = V[:, -1]
    ph = v[0] / abs(v[0]) if abs(v[0]) > 0 else 1.0
    return v / ph


def ecalib(ksp, calsize=(24, 24, 24), kdims=(6, 6, 6), threshold=0.001, crop=0.8, maps=2, iters=30,
           rotphase=True):
    """ESPIRiT (src/ecalib.c:47-263 with calib2, calib/calib.c:715-798):
    returns (maps [M, C, Z, Y, X], eigenvalue maps [M, Z, Y, X], number of
    kernels)."""
    ksp = np.asarray(ksp, np.complex128)
    C, Z, Y, X = ksp.shape
    kd = [k if n > 1 else 1 for k, n in zip(kdims, (X, Y, Z))]
    cal = calib_geom(ksp, calsize)
    G = casorati_gram(cal, kd)
    kerns, val, nk = compute_kernels(G, threshold)
    Gr = compute_imgcov(kerns, kd, C)
    big = sinc_zeropad(Gr, (Z, Y, X))
    P = Z * Y * X
    Hs = big.reshape(P, C, C)
    Hs = np.tril(Hs) + np.conj(np.transpose(np.tril(Hs, -1), (0, 2, 1)))
    rows, vals = orthiter_batch(maps, Hs, iters)
    sens = np.zeros((maps, C, P), np.complex128)
    ev = np.zeros((maps, P))
    for u in range(maps):
        sens[u] = rows[:, maps - 1 - u, :].T
        ev[u] = vals[:, maps - 1 - u]
    sens = sens.reshape(maps, C, Z, Y, X)
    ev = ev.reshape(maps, Z, Y, X)
    # crop_sens (:161-165): maps where |eigenvalue| >= crop
    sens = sens * (np.abs(ev) >= np.float32(crop))[:, None]
    if rotphase:
        r = scc_rot(cal)
        t = np.einsum("uczyx,c->uzyx", sens, np.conj(r))
        ph = np.where(np.abs(t) > 0, t / np.where(np.abs(t) > 0, np.abs(t), 1.0), 1.0)
        sens = sens * np.conj(ph)[:, None]
    return sens, ev, nk
