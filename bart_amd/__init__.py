"""bart_amd -- gfx950 NUFFT / CG-SENSE behind BART's linop interface.

Python mirror of the C ABI in ``include/bart_amd.h`` (``libbart_amd.so``):
``nufft_create2`` / ``nufft_create`` / ``sense_nc_init`` build operators,
``linop_forward`` / ``linop_adjoint`` / ``linop_normal`` apply them,
``lsqr_conjgrad`` runs CG on the normal equations, ``fft`` / ``ifft`` are the
batched FFTs.  Arrays use BART's layout: a list of (normally 16) dims,
column-major, complex64.  Arguments may be numpy arrays (host) or torch CUDA
tensors (device); the library selects the device path by pointer location as
BART does.

There is no CPU fallback: if the HIP library is missing or no GPU is present
the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbart_amd.so")
_LIB = None

DIMS = 16


class BartError(RuntimeError):
    pass


class NufftConf(ctypes.Structure):
    """struct nufft_conf_s (noncart/nufft.h:7-29)."""
    _fields_ = [
        ("toeplitz", ctypes.c_bool),
        ("pcycle", ctypes.c_bool),
        ("periodic", ctypes.c_bool),
        ("lowmem", ctypes.c_bool),
        ("flags", ctypes.c_ulong),
        ("cfft", ctypes.c_ulong),
        ("decomp", ctypes.c_bool),
        ("nopsf", ctypes.c_bool),
        ("upper_triag", ctypes.c_bool),
        ("real", ctypes.c_bool),
        ("compress_psf", ctypes.c_bool),
        ("decomposed_psf", ctypes.c_bool),
        ("precomp", ctypes.c_bool),
        ("precomp_linphase", ctypes.c_bool),
        ("precomp_fftmod", ctypes.c_bool),
        ("precomp_roll", ctypes.c_bool),
        ("zero_overhead", ctypes.c_bool),
        ("width", ctypes.c_float),
        ("os", ctypes.c_float),
    ]

    @classmethod
    def defaults(cls, **kw):
        """nufft_conf_defaults (noncart/nufft.c:53-75) with overrides."""
        c = cls(toeplitz=True, pcycle=False, periodic=False, lowmem=False, flags=7, cfft=0,
                decomp=True, nopsf=False, upper_triag=False, real=False, compress_psf=False,
                decomposed_psf=False, precomp=True, precomp_linphase=True, precomp_fftmod=True,
                precomp_roll=True, zero_overhead=False, width=6.0, os=2.0)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class CgStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("status", ctypes.c_int),
        ("resid", ctypes.c_float),
        ("bnorm", ctypes.c_float),
        ("time_ms", ctypes.c_float),
        ("fell_back", ctypes.c_bool),
    ]


def library_path():
    return _LIB_PATH


def build(verbose=False):
    """Compile libbart_amd.so for gfx950 (hipcc, in-tree)."""
    import subprocess
    out = None if verbose else subprocess.DEVNULL
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-j", jobs, "-C", os.path.join(_HERE, "csrc")], check=True, stdout=out)


def _init_torch_first():
    """PyTorch-ROCm wheels bundle their own HIP runtime.  When torch is used in
    the same process (as device-memory plumbing), its runtime must initialise
    the GPU before this library's does, or torch's device init fails.  Only
    done when torch is already imported; a device count query does not
    initialise the GPU."""
    import sys
    torch = sys.modules.get("torch")
    if torch is None:
        return
    try:
        if torch.cuda.device_count() > 0:
            torch.cuda.init()
    except Exception:
        pass


def lib():
    """Load libbart_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(_LIB_PATH):
        raise BartError("libbart_amd.so not built: run bart_amd.build() (hipcc, gfx950)")
    _init_torch_first()
    L = ctypes.CDLL(_LIB_PATH)
    vp = ctypes.c_void_p
    lp = ctypes.POINTER(ctypes.c_long)
    L.nufft_create2.restype = vp
    L.nufft_create2.argtypes = [ctypes.c_int, lp, lp, lp, vp, lp, vp, lp, vp, NufftConf]
    L.nufft_create.restype = vp
    L.nufft_create.argtypes = [ctypes.c_int, lp, lp, lp, vp, vp, NufftConf]
    L.sense_nc_init.restype = vp
    L.sense_nc_init.argtypes = [lp, lp, vp, lp, lp, vp, ctypes.POINTER(NufftConf), lp, vp, lp, vp, vp, ctypes.c_ulong]
    L.nudft_create.restype = vp
    L.nudft_create.argtypes = [ctypes.c_int, ctypes.c_ulong, lp, lp, lp, vp, lp, vp, lp, vp]
    for f in (L.linop_forward, L.linop_adjoint):
        f.restype = None
        f.argtypes = [vp, ctypes.c_int, lp, vp, ctypes.c_int, lp, vp]
    L.linop_normal.restype = None
    L.linop_normal.argtypes = [vp, ctypes.c_int, lp, vp, vp]
    L.linop_free.restype = None
    L.linop_free.argtypes = [vp]
    L.nufft_get_psf_dims.restype = ctypes.c_int
    L.nufft_get_psf_dims.argtypes = [vp, ctypes.c_int, lp]
    L.nufft_get_psf.restype = None
    L.nufft_get_psf.argtypes = [vp, ctypes.c_int, lp, vp]
    L.nufft_update_psf.restype = None
    L.nufft_update_psf.argtypes = [vp, ctypes.c_int, lp, vp]
    L.nufft_get_psf2.restype = None
    L.nufft_get_psf2.argtypes = [vp, ctypes.c_int, lp, lp, vp]
    L.nufft_update_psf2.restype = None
    L.nufft_update_psf2.argtypes = [vp, ctypes.c_int, lp, lp, vp]
    L.nufft_create_normal.restype = vp
    L.nufft_create_normal.argtypes = [ctypes.c_int, lp, ctypes.c_int, lp, vp, ctypes.c_bool, NufftConf]
    L.nufft_update_traj.restype = None
    L.nufft_update_traj.argtypes = [vp, ctypes.c_int, lp, vp, lp, vp, lp, vp]
    L.compute_psf2.restype = vp
    L.compute_psf2.argtypes = [ctypes.c_int, lp, ctypes.c_ulong, lp, vp, lp, vp, lp, vp, ctypes.c_bool,
                               ctypes.c_bool, ctypes.c_bool]
    L.bart_amd_linop_set_stream.restype = ctypes.c_int
    L.bart_amd_linop_set_stream.argtypes = [vp, vp]
    for f in (L.fft, L.ifft):
        f.restype = None
        f.argtypes = [ctypes.c_int, lp, ctypes.c_ulong, vp, vp]
    L.lsqr_conjgrad.restype = ctypes.c_int
    L.lsqr_conjgrad.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, lp, vp, vp,
                                ctypes.c_bool, ctypes.c_bool, vp, ctypes.POINTER(CgStats)]
    L.bart_amd_warmnet_input.restype = ctypes.c_int
    L.bart_amd_warmnet_input.argtypes = [vp, ctypes.c_float, ctypes.c_int, lp, vp, vp, ctypes.POINTER(ctypes.c_float)]
    L.bart_amd_warmnet_pack.restype = ctypes.c_int
    L.bart_amd_warmnet_pack.argtypes = [ctypes.c_int, lp, vp, ctypes.c_float, vp]
    L.bart_amd_warmnet_guess.restype = ctypes.c_int
    L.bart_amd_warmnet_guess.argtypes = [vp, vp, ctypes.c_float, ctypes.c_int, lp, vp, vp]
    L.bart_amd_warmnet_input_dir.restype = ctypes.c_int
    L.bart_amd_warmnet_input_dir.argtypes = [vp, ctypes.c_float, ctypes.c_int, lp, vp, vp, vp,
                                             ctypes.POINTER(ctypes.c_float)]
    L.bart_amd_warmnet_guess_dir.restype = ctypes.c_int
    L.bart_amd_warmnet_guess_dir.argtypes = [vp, vp, ctypes.c_float, ctypes.c_int, lp, vp, vp, vp]
    L.bart_amd_warmnet_guess_dcf.restype = ctypes.c_int
    L.bart_amd_warmnet_guess_dcf.argtypes = [vp, vp, ctypes.c_float, ctypes.c_int, lp, vp, vp, vp, ctypes.c_long, vp]
    L.nufft_precond_create.restype = vp
    L.nufft_precond_create.argtypes = [vp]
    L.operator_free.restype = None
    L.operator_free.argtypes = [vp]
    L.bart_amd_precond_apply.restype = ctypes.c_int
    L.bart_amd_precond_apply.argtypes = [vp, vp, vp]
    L.lsqr_conjgrad_precond.restype = ctypes.c_int
    L.lsqr_conjgrad_precond.argtypes = [vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, lp, vp,
                                        vp, ctypes.c_bool, vp, ctypes.POINTER(CgStats)]
    L.bart_amd_tuning_report.restype = ctypes.c_int
    L.bart_amd_tuning_report.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.bart_amd_cg_monitor.restype = ctypes.c_int
    L.bart_amd_cg_monitor.argtypes = [vp, vp, vp]
    L.lsqr_conjgrad_noalloc.restype = ctypes.c_int
    L.lsqr_conjgrad_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, lp, vp, vp,
                                        ctypes.c_bool, ctypes.c_bool, vp, ctypes.POINTER(CgStats)]
    L.lsqr_conjgrad_batch.restype = ctypes.c_int
    L.lsqr_conjgrad_batch.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_bool,
                                      ctypes.c_bool, vp, ctypes.POINTER(CgStats)]
    ci = ctypes.c_int
    L.bart_amd_conv3x3.restype = ci
    L.bart_amd_conv3x3.argtypes = [vp, ci, ci, ci, ci, vp, vp, vp, ci, ci, vp]
    L.bart_amd_conv3x3_bwd_data.restype = ci
    L.bart_amd_conv3x3_bwd_data.argtypes = [vp, ci, ci, ci, ci, vp, ci, vp]
    L.bart_amd_conv3x3_bwd_weights.restype = ci
    L.bart_amd_conv3x3_bwd_weights.argtypes = [vp, vp, ci, ci, ci, ci, ci, vp, vp]
    L.bart_amd_warmnet_create.restype = vp
    L.bart_amd_warmnet_create.argtypes = [ci, ctypes.c_ulong]
    L.bart_amd_warmnet_free.restype = None
    L.bart_amd_warmnet_free.argtypes = [vp]
    L.bart_amd_warmnet_num_params.restype = ctypes.c_long
    L.bart_amd_warmnet_num_params.argtypes = [vp]
    L.bart_amd_warmnet_params.restype = ci
    L.bart_amd_warmnet_params.argtypes = [vp, vp, ci]
    L.bart_amd_warmnet_forward.restype = ci
    L.bart_amd_warmnet_forward.argtypes = [vp, vp, ci, ci, ci, vp]
    L.bart_amd_warmnet_train_step.restype = ctypes.c_double
    L.bart_amd_warmnet_train_step.argtypes = [vp, vp, vp, ci, ci, ci, ctypes.c_float]
    L.bart_amd_warmnet_train_step_grad.restype = ci
    L.bart_amd_warmnet_train_step_grad.argtypes = [vp, vp, vp, ci, ci, ci, ctypes.c_float]
    L.extract_calib.restype = vp
    L.extract_calib.argtypes = [vp, vp, vp, vp, ctypes.c_bool]
    L.calib.restype = None
    L.calib.argtypes = [vp, vp, vp, vp, ci, vp, vp, vp]
    L.bart_amd_set_error_mode.argtypes = [ctypes.c_int]
    L.bart_amd_last_error.restype = ctypes.c_char_p
    L.bart_amd_clear_error.restype = None
    L.bart_amd_version.restype = ctypes.c_char_p
    L.bart_amd_set_error_mode(1)   # raise in Python instead of exit(1)
    _LIB = L
    return L


def _check():
    e = lib().bart_amd_last_error()
    if e:
        lib().bart_amd_clear_error()
        raise BartError(e.decode())


def _dims(d, n=DIMS):
    d = [int(v) for v in d] + [1] * (n - len(d))
    return (ctypes.c_long * n)(*d[:n])


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _ptr(a):
    """Pointer of a contiguous complex64 numpy array or torch tensor."""
    if a is None:
        return None
    if _is_torch(a):
        assert a.is_contiguous() and str(a.dtype) == "torch.complex64", "need contiguous complex64 tensor"
        return ctypes.c_void_p(a.data_ptr())
    assert a.dtype == np.complex64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(ctypes.c_void_p)


def _like(src, shape):
    """Output array on the same side (host/device) as src, BART dims shape."""
    n = int(np.prod(shape))
    if _is_torch(src):
        import torch
        return torch.empty(n, dtype=torch.complex64, device=src.device)
    return np.empty(n, np.complex64)


def _host_c64(a):
    """complex64 array whose memory is taken as BART's column-major layout:
    a C-ordered array of shape dims[::-1] or an F-ordered array of shape dims
    (outputs are flat)."""
    if a is None or _is_torch(a):
        return a
    a = np.asarray(a, np.complex64)
    if not (a.flags.c_contiguous or a.flags.f_contiguous):
        a = np.ascontiguousarray(a)
    return a


def _torch_stream(*arrays):
    """torch's current stream for the first CUDA tensor among arrays (0 =
    the default stream, on which the library's own blocking stream is
    ordered implicitly), so that the library runs after the work that wrote
    its inputs and before the work that reads its outputs."""
    for a in arrays:
        if a is not None and _is_torch(a):
            import torch
            return int(torch.cuda.current_stream(a.device).cuda_stream)
    return 0


class Linop:
    """A BART linop_s handle with its output (odims) and input (idims) dims."""

    def __init__(self, handle, odims, idims):
        if not handle:
            _check()
            raise BartError("operator creation failed")
        self.h = handle
        self.odims = list(odims) + [1] * (DIMS - len(odims))
        self.idims = list(idims) + [1] * (DIMS - len(idims))

    def free(self):
        if getattr(self, "h", None):
            lib().linop_free(self.h)
            self.h = None
            _check()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def set_stream(self, stream):
        """Launch this handle's applies on a hipStream_t (int handle; 0: its
        own stream)."""
        lib().bart_amd_linop_set_stream(self.h, ctypes.c_void_p(int(stream)))
        _check()

    def _stream_of(self, *arrays):
        st = _torch_stream(*arrays)
        if st != getattr(self, "_stream", 0):
            self.set_stream(st)
            self._stream = st

    def forward(self, x, out=None):
        x = _host_c64(x)
        out = _like(x, self.odims) if out is None else out
        self._stream_of(x, out)
        lib().linop_forward(self.h, DIMS, _dims(self.odims), _ptr(out), DIMS, _dims(self.idims), _ptr(x))
        _check()
        return out

    def adjoint(self, y, out=None):
        y = _host_c64(y)
        out = _like(y, self.idims) if out is None else out
        self._stream_of(y, out)
        lib().linop_adjoint(self.h, DIMS, _dims(self.idims), _ptr(out), DIMS, _dims(self.odims), _ptr(y))
        _check()
        return out

    def normal(self, x, out=None):
        x = _host_c64(x)
        out = _like(x, self.idims) if out is None else out
        self._stream_of(x, out)
        lib().linop_normal(self.h, DIMS, _dims(self.idims), _ptr(out), _ptr(x))
        _check()
        return out

    def psf_dims(self):
        n = lib().nufft_get_psf_dims(self.h, 0, None)
        _check()
        d = (ctypes.c_long * n)()
        lib().nufft_get_psf_dims(self.h, n, d)
        _check()
        return list(d)

    def get_psf(self):
        d = self.psf_dims()
        out = np.empty(int(np.prod(d)), np.complex64)
        lib().nufft_get_psf(self.h, len(d), _dims(d, len(d)), _ptr(out))
        _check()
        return out

    def update_psf(self, psf):
        d = self.psf_dims()
        psf = _host_c64(psf)
        lib().nufft_update_psf(self.h, len(d), _dims(d, len(d)), _ptr(psf))
        _check()

    def get_psf2(self, out, strides):
        """nufft_get_psf2: the PSF into out (complex64, host or device) at
        the given byte strides."""
        d = self.psf_dims()
        lib().nufft_get_psf2(self.h, len(d), _dims(d, len(d)), _dims(strides, len(d)), _ptr(out))
        _check()
        return out

    def update_psf2(self, psf, strides):
        d = self.psf_dims()
        lib().nufft_update_psf2(self.h, len(d), _dims(d, len(d)), _dims(strides, len(d)), _ptr(psf))
        _check()

    def update_traj(self, traj_dims, traj, weights=None, wgh_dims=None):
        """nufft_update_traj: a new trajectory (and weights) of the same
        sample count; the plan and, in Toeplitz mode, the PSF are rebuilt."""
        traj = _host_c64(traj) if not _is_torch(traj) else traj
        w = _host_c64(weights) if weights is not None and not _is_torch(weights) else weights
        if weights is not None and wgh_dims is None:
            wgh_dims = [1] + list(traj_dims[1:3])
        lib().nufft_update_traj(self.h, DIMS, _dims(traj_dims), _ptr(traj),
                                _dims(wgh_dims) if wgh_dims is not None else None, _ptr(w), None, None)
        _check()


def traj_array(traj_m3):
    """[M, 3] coordinates -> BART trajectory memory [3, M] complex64."""
    t = np.asarray(traj_m3)
    return np.ascontiguousarray(t.reshape(-1, 3).astype(np.complex64))


def nufft_create2(ksp_dims, cim_dims, traj_dims, traj, wgh_dims=None, weights=None, conf=None, bas_dims=None,
                  basis=None):
    """noncart/nufft.h:48-57.  With a subspace basis (bas_dims [1 1 1 1 1 T
    K]): k-space [1 R S C 1 T], coil images [X Y Z C 1 1 K]."""
    conf = NufftConf.defaults() if conf is None else conf
    traj = _host_c64(traj) if not _is_torch(traj) else traj
    w = _host_c64(weights) if weights is not None and not _is_torch(weights) else weights
    bas = _host_c64(basis) if basis is not None and not _is_torch(basis) else basis
    if weights is not None and wgh_dims is None:
        wgh_dims = [1] + list(traj_dims[1:])
    h = lib().nufft_create2(DIMS, _dims(ksp_dims), _dims(cim_dims), _dims(traj_dims), _ptr(traj),
                            _dims(wgh_dims) if wgh_dims is not None else None, _ptr(w),
                            _dims(bas_dims) if bas_dims is not None else None, _ptr(bas), conf)
    odims = list(ksp_dims) + [1] * (DIMS - len(ksp_dims))
    if basis is not None:
        odims[6] = 1
    return Linop(h, odims, cim_dims)


def nufft_create(ksp_dims, cim_dims, traj_dims, traj, weights=None, conf=None):
    """noncart/nufft.h:40-46."""
    conf = NufftConf.defaults() if conf is None else conf
    traj = _host_c64(traj) if not _is_torch(traj) else traj
    w = _host_c64(weights) if weights is not None and not _is_torch(weights) else weights
    h = lib().nufft_create(DIMS, _dims(ksp_dims), _dims(cim_dims), _dims(traj_dims), _ptr(traj), _ptr(w), conf)
    return Linop(h, ksp_dims, cim_dims)


def nufft_create_normal(cim_dims, psf_dims, psf=None, conf=None):
    """noncart/nufft.h:84-86: the Toeplitz normal operator of a PSF alone
    (coil images -> coil images; forward and adjoint are errors)."""
    conf = NufftConf.defaults() if conf is None else conf
    psf = _host_c64(psf) if psf is not None and not _is_torch(psf) else psf
    h = lib().nufft_create_normal(DIMS, _dims(cim_dims), DIMS + 1, _dims(psf_dims, DIMS + 1), _ptr(psf), False,
                                  conf)
    return Linop(h, [1] * DIMS, cim_dims)


def compute_psf2(psf_dims, traj_dims, traj, weights=None, wgh_dims=None, periodic=True):
    """noncart/nufft.h:70-73: the decomposed PSF of a trajectory
    (psf_dims: DIMS + 1 dims, image dims and 2^d last) as a host array."""
    traj = _host_c64(traj)
    w = _host_c64(weights) if weights is not None else None
    if weights is not None and wgh_dims is None:
        wgh_dims = [1] + list(traj_dims[1:3])
    n = int(np.prod(psf_dims))
    p = lib().compute_psf2(DIMS, _dims(psf_dims, DIMS + 1), 7, _dims(traj_dims, DIMS + 1), _ptr(traj), None, None,
                           _dims(wgh_dims, DIMS + 1) if wgh_dims is not None else None, _ptr(w), periodic, False,
                           False)
    _check()
    if not p:
        raise BartError("compute_psf2 failed")
    out = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_float)), shape=(2 * n,)).copy()
    ctypes.CDLL(None).free(ctypes.c_void_p(p))
    return out.view(np.complex64)


def nudft_create(ksp_dims, cim_dims, traj_dims, traj):
    """noncart/nudft.c:283-348: the direct non-uniform DFT (nufft -s)."""
    traj = _host_c64(traj) if not _is_torch(traj) else traj
    h = lib().nudft_create(DIMS, 7, _dims(ksp_dims), _dims(cim_dims), _dims(traj_dims), _ptr(traj), None, None,
                           None, None)
    return Linop(h, ksp_dims, cim_dims)


def sense_nc_init(max_dims, map_dims, maps, ksp_dims, traj_dims, traj, conf=None, wgs_dims=None, weights=None,
                  nufft_op=False, basis_dims=None, basis=None):
    """sense/modelnc.c:28-67: A = NUFFT S, image -> k-space.  With nufft_op
    the NUFFT operator sharing the plan (BART's fft_opp) is returned too.
    With a subspace basis (basis_dims [1 1 1 1 1 T K]): max_dims dim 6 = K,
    k-space [1 R S C 1 T]."""
    conf = NufftConf.defaults() if conf is None else conf
    maps = _host_c64(maps) if not _is_torch(maps) else maps
    traj = _host_c64(traj) if not _is_torch(traj) else traj
    w = _host_c64(weights) if weights is not None and not _is_torch(weights) else weights
    bas = _host_c64(basis) if basis is not None and not _is_torch(basis) else basis
    if weights is not None and wgs_dims is None:
        wgs_dims = [1] + list(traj_dims[1:3]) + [1] * 13
    fop = ctypes.c_void_p(0)
    h = lib().sense_nc_init(_dims(max_dims), _dims(map_dims), _ptr(maps), _dims(ksp_dims), _dims(traj_dims),
                            _ptr(traj), ctypes.byref(conf), _dims(wgs_dims) if wgs_dims is not None else None,
                            _ptr(w), _dims(basis_dims) if basis_dims is not None else None, _ptr(bas),
                            ctypes.byref(fop) if nufft_op else None, 0)
    img_dims = list(max_dims) + [1] * (DIMS - len(max_dims))
    img_dims[3] = 1
    img_dims[4] = 1
    op = Linop(h, ksp_dims, img_dims)
    if not nufft_op:
        return op
    cim_dims = list(max_dims) + [1] * (DIMS - len(max_dims))
    cim_dims[4] = 1
    return op, Linop(fop.value, ksp_dims, cim_dims)


def lsqr_conjgrad(op, b, maxiter, lam=0.0, tol=0.0, x0=None, fallback=False):
    """CG on (A^H A + lam I) x = b.  Returns (x, stats, residual trace)."""
    b = _host_c64(b)
    warm = x0 is not None
    if warm:
        x = x0.clone() if _is_torch(x0) else np.array(_host_c64(x0), copy=True)
    else:
        x = _like(b, op.idims)
    trace = np.zeros(max(1, maxiter), np.float32)
    st = CgStats()
    it = lib().lsqr_conjgrad(op.h, maxiter, lam, tol, DIMS, _dims(op.idims), _ptr(x), _ptr(b), warm, fallback,
                             trace.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st))
    _check()
    return x, st, trace[:max(0, it)]


class Precond:
    """Strang's circulant preconditioner of a Toeplitz operator
    (nufft_precond_create)."""

    def __init__(self, op):
        self.op = op
        self.h = lib().nufft_precond_create(op.h)
        _check()

    def apply(self, x):
        x = _host_c64(x)
        out = _like(x, self.op.idims)
        lib().bart_amd_precond_apply(self.h, _ptr(out), _ptr(x))
        _check()
        return out

    def free(self):
        if getattr(self, "h", None):
            lib().operator_free(self.h)
            self.h = None


class EcalibConf(ctypes.Structure):
    """struct ecalib_conf (calib/calib.h:12-34)."""
    _fields_ = [
        ("kdims", ctypes.c_long * 3),
        ("threshold", ctypes.c_float),
        ("numsv", ctypes.c_int),
        ("percentsv", ctypes.c_float),
        ("weighting", ctypes.c_bool),
        ("softcrop", ctypes.c_bool),
        ("crop", ctypes.c_float),
        ("orthiter", ctypes.c_bool),
        ("num_orthiter", ctypes.c_int),
        ("usegpu", ctypes.c_bool),
        ("perturb", ctypes.c_float),
        ("intensity", ctypes.c_bool),
        ("rotphase", ctypes.c_bool),
        ("var", ctypes.c_float),
        ("automate", ctypes.c_bool),
        ("phase_normalize", ctypes.c_bool),
        ("econdim", ctypes.c_int),
        ("nystroem", ctypes.c_bool),
        ("nystroem_os", ctypes.c_long),
        ("nystroem_K", ctypes.c_long),
    ]

    @classmethod
    def defaults(cls, **kw):
        """ecalib_defaults (calib/calib.c:686-708) with overrides."""
        c = cls()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(cls.in_dll(lib(), "ecalib_defaults")), ctypes.sizeof(cls))
        for k, v in kw.items():
            if "kdims" == k:
                c.kdims = (ctypes.c_long * 3)(*v)
            else:
                setattr(c, k, v)
        return c


def ecalib(ksp, maps=2, calsize=(24, 24, 24), conf=None):
    """ESPIRiT (src/ecalib.c:47-263): Cartesian k-space [C, Z, Y, X] ->
    (maps [M, C, Z, Y, X], eigenvalue maps [M, Z, Y, X], singular values),
    via extract_calib and calib of the C ABI."""
    conf = EcalibConf.defaults() if conf is None else conf
    ksp = np.ascontiguousarray(ksp, np.complex64)
    C, Z, Y, X = ksp.shape
    kd = _dims([X, Y, Z, C])
    cd = (ctypes.c_long * DIMS)()
    cs = (ctypes.c_long * 3)(*calsize)
    p = lib().extract_calib(cd, cs, kd, _ptr(ksp), False)
    _check()
    if not p:
        raise BartError("extract_calib failed")
    try:
        for i in range(3):
            if cd[i] == 1 and [X, Y, Z][i] != 1:
                raise BartError("Calibration region not found!")
        kdims = [conf.kdims[i] if [X, Y, Z][i] > 1 else 1 for i in range(3)]
        conf.kdims = (ctypes.c_long * 3)(*kdims)
        od = [X if kdims[0] > 1 else 1, Y if kdims[1] > 1 else 1, Z if kdims[2] > 1 else 1, C, maps]
        out = np.zeros([maps, C] + od[2::-1], np.complex64)
        ev = np.zeros([maps] + od[2::-1], np.complex64)
        N = int(np.prod(kdims)) * C
        sv = np.zeros(N, np.float32)
        lib().calib(ctypes.byref(conf), _dims(od), _ptr(out), _ptr(ev), N,
                    sv.ctypes.data_as(ctypes.c_void_p), cd, ctypes.c_void_p(p))
        _check()
    finally:
        ctypes.CDLL(None).free(ctypes.c_void_p(p))
    return out, ev.real.copy(), sv


def lsqr_conjgrad_precond(op, pre, b, maxiter, lam=0.0, tol=0.0):
    """CG on P (A^H A + lam I) x = P b (nufft -i -c).  Returns (x, stats,
    residual trace)."""
    b = _host_c64(b)
    x = _like(b, op.idims)
    trace = np.zeros(max(1, maxiter), np.float32)
    st = CgStats()
    it = lib().lsqr_conjgrad_precond(op.h, pre.h, maxiter, lam, tol, DIMS, _dims(op.idims), _ptr(x), _ptr(b), False,
                                     trace.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st))
    _check()
    return x, st, trace[:max(0, it)]


def tuning_report():
    """The autotuner's choices in this process (list of dicts: shape class,
    coils per work-group, work-groups per compute unit, ms, candidates)."""
    import json
    n = lib().bart_amd_tuning_report(None, 0)
    buf = ctypes.create_string_buffer(n + 16)
    lib().bart_amd_tuning_report(buf, n + 16)
    _check()
    return json.loads(buf.value.decode())


class CgMonitor:
    """Truth-image monitor of op's CG solves (pics -T): after each
    lsqr_conjgrad on op, errors()[k] = ||truth - x_k|| / ||truth|| for the
    iterate x_k that iteration k started from."""

    def __init__(self, op, truth, maxiter):
        self.op = op
        self.err = np.zeros(max(1, maxiter), np.float32)
        lib().bart_amd_cg_monitor(op.h, _ptr(_host_c64(truth)), self.err.ctypes.data_as(ctypes.c_void_p))
        _check()

    def errors(self, n):
        return self.err[:n].copy()

    def close(self):
        lib().bart_amd_cg_monitor(self.op.h, None, None)
        _check()


def lsqr_conjgrad_noalloc(op, b, maxiter, lam=0.0, tol=0.0, x0=None, fallback=False):
    """lsqr_conjgrad in the library's static working memory (no device
    allocation by the library): b and x0 must be CUDA tensors.  Returns
    (x, stats, residual trace)."""
    assert _is_torch(b) and b.is_cuda, "lsqr_conjgrad_noalloc takes device tensors"
    warm = x0 is not None
    x = x0.clone() if warm else _like(b, op.idims)
    import torch
    torch.cuda.synchronize()
    trace = np.zeros(max(1, maxiter), np.float32)
    st = CgStats()
    it = lib().lsqr_conjgrad_noalloc(op.h, maxiter, lam, tol, DIMS, _dims(op.idims), _ptr(x), _ptr(b), warm,
                                     fallback, trace.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st))
    _check()
    return x, st, trace[:max(0, it)]


def lsqr_conjgrad_batch(ops, bs, maxiter, lam=0.0, tol=0.0, x0s=None, fallback=False):
    """Independent CG solves (A_i^H A_i + lam I) x_i = bs[i], one operator
    each, run together on the device.  Returns (xs, stats, traces)."""
    n = len(ops)
    assert len(bs) == n and (x0s is None or len(x0s) == n)
    bs = [_host_c64(b) for b in bs]
    warm = x0s is not None
    if warm:
        xs = [x.clone() if _is_torch(x) else np.array(_host_c64(x), copy=True) for x in x0s]
    else:
        xs = [_like(b, op.idims) for b, op in zip(bs, ops)]
    T = max(1, maxiter)
    trace = np.zeros((max(1, n), T), np.float32)
    stats = (CgStats * max(1, n))()
    vp = ctypes.c_void_p
    hs = (vp * max(1, n))(*[op.h for op in ops])
    xp = (vp * max(1, n))(*[_ptr(x).value for x in xs])
    bp = (vp * max(1, n))(*[_ptr(b).value for b in bs])
    lib().lsqr_conjgrad_batch(n, hs, maxiter, lam, tol, DIMS, xp, bp, warm, fallback,
                              trace.ctypes.data_as(vp), stats)
    _check()
    st = [stats[i] for i in range(n)]
    return xs, st, [trace[i, :max(0, st[i].iterations)] for i in range(n)]


def _fptr(t):
    """Device pointer of a contiguous float32 torch tensor (or None)."""
    if t is None:
        return None
    assert _is_torch(t) and t.is_cuda and t.is_contiguous() and str(t.dtype) == "torch.float32", \
        "need a contiguous float32 CUDA tensor"
    return ctypes.c_void_p(t.data_ptr())


def conv3x3(x, w, bias=None, res=None, relu=False):
    """Warm-start network layer: [relu](conv3x3(x, w) + bias) [+ res] on the
    matrix cores (bfloat16 operands, FP32 accumulation).  x [B, Cin, H, W],
    w [Cout, Cin, 3, 3], float32 CUDA tensors."""
    import torch
    B, Cin, H, W = x.shape
    Cout = w.shape[0]
    y = torch.empty((B, Cout, H, W), dtype=torch.float32, device=x.device)
    lib().bart_amd_conv3x3(_fptr(x), B, Cin, H, W, _fptr(w), _fptr(bias), _fptr(res), Cout, 1 if relu else 0,
                           _fptr(y))
    _check()
    return y


def conv3x3_bwd_data(dy, w):
    """Gradient of conv3x3 with respect to its input for output gradient dy."""
    import torch
    B, Cout, H, W = dy.shape
    Cin = w.shape[1]
    dx = torch.empty((B, Cin, H, W), dtype=torch.float32, device=dy.device)
    lib().bart_amd_conv3x3_bwd_data(_fptr(dy), B, Cout, H, W, _fptr(w), Cin, _fptr(dx))
    _check()
    return dx


def conv3x3_bwd_weights(x, dy, bias=True):
    """Gradients of conv3x3 with respect to w [Cout, Cin, 3, 3] and bias."""
    import torch
    B, Cin, H, W = x.shape
    Cout = dy.shape[1]
    dw = torch.empty((Cout, Cin, 3, 3), dtype=torch.float32, device=x.device)
    db = torch.empty(Cout, dtype=torch.float32, device=x.device) if bias else None
    lib().bart_amd_conv3x3_bwd_weights(_fptr(x), _fptr(dy), B, Cin, Cout, H, W, _fptr(dw), _fptr(db))
    _check()
    return dw, db


def fft(dims, flags, x, inverse=False):
    """num/fft.c:387-417 (unnormalised, -1 forward / +1 inverse)."""
    x = _host_c64(x)
    out = _like(x, dims)
    f = lib().ifft if inverse else lib().fft
    f(len(dims), _dims(dims, len(dims)), flags, _ptr(out), _ptr(x))
    _check()
    return out


def ifft(dims, flags, x):
    return fft(dims, flags, x, inverse=True)


def version():
    return lib().bart_amd_version().decode()
