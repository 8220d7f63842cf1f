"""CPU checks of the C ABI: the library builds, loads without a GPU and
exports every function include/bart_amd.h declares; the conf struct layout
matches the header; the host-side mirror fails loudly without the library."""
import ctypes
import os
import re

import pytest

import bart_amd as ba

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "bart_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof")))


def test_header_declares_the_boundary():
    fns = header_functions()
    for f in ["nufft_create2", "nufft_create", "sense_nc_init", "linop_forward", "linop_adjoint", "linop_normal",
              "linop_free", "nufft_get_psf_dims", "nufft_get_psf", "nufft_update_psf", "fft", "ifft",
              "lsqr_conjgrad", "nufft_get_psf2", "nufft_update_psf2", "nufft_create_normal", "nufft_update_traj",
              "compute_psf2", "bart_amd_linop_set_stream"]:
        assert f in fns


def test_library_exports_every_declared_symbol():
    if not os.path.exists(ba.library_path()):
        ba.build()
    L = ctypes.CDLL(ba.library_path())
    missing = [f for f in header_functions() if not hasattr(L, f)]
    assert not missing, missing
    # data symbol
    assert ctypes.c_bool.in_dll(L, "nufft_conf_defaults").value is True


def test_conf_defaults_layout():
    L = ctypes.CDLL(ba.library_path())
    c = ba.NufftConf.in_dll(L, "nufft_conf_defaults")
    ref = ba.NufftConf.defaults()
    for name, _ in ba.NufftConf._fields_:
        assert getattr(c, name) == getattr(ref, name), name
    assert c.width == 6.0 and c.os == 2.0 and c.flags == 7


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(ba, "_LIB", None)
    monkeypatch.setattr(ba, "_LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ba.BartError):
        ba.lib()


def test_basis_dims_are_checked_before_any_device_call():
    # nufft_create2 with a subspace basis rejects dims that do not fit
    # (image dim 6 = coefficients, k-space dim 5 = the basis frames)
    import numpy as np
    from bart_amd import inputs as I
    d16 = lambda *d: list(d) + [1] * (16 - len(d))
    n, Mf = 16, 128
    traj = np.ascontiguousarray(I.radial_traj(32, 8).reshape(2, Mf, 3))
    B = np.ones((2, 2), np.complex64)
    conf = ba.NufftConf.defaults()
    with pytest.raises(ba.BartError, match="coefficients"):
        ba.nufft_create2(d16(1, Mf, 1, 1, 1, 2), d16(n, n, 1, 1), d16(3, Mf, 1, 1, 1, 2), traj, conf=conf,
                         bas_dims=d16(1, 1, 1, 1, 1, 2, 2), basis=B)
    with pytest.raises(ba.BartError, match="dim 5"):
        ba.nufft_create2(d16(1, Mf, 1, 1, 1, 3), d16(n, n, 1, 1, 1, 1, 2), d16(3, Mf, 1, 1, 1, 2), traj,
                         conf=conf, bas_dims=d16(1, 1, 1, 1, 1, 2, 2), basis=B)
