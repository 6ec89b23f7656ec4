"""THC ERI file (thc.eri.h5) reader and writer: ctypes binding of libthcgw_eri.so
(include/thcgw_eri.h).

The file is the one the reference's thc_reader_t writes after an ISDF build and reads back
(src/methods/ERI/thc_reader_t.hpp:236-256, 507-574).  read_thc_eri returns host arrays
in the layouts of the GPU path: X_skPa (ns*npol, nk, Np, nx) and V_qPQ (nq_ibz, Np, Np),
the latter being coulomb_matrix as stored (thc_reader_t::Z, N_k factor included).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ERI_LIB_PATH = os.path.join(_HERE, "libthcgw_eri.so")
ERI_EXPORTS = ("thcgw_eri_read_info", "thcgw_eri_read_X", "thcgw_eri_read_V", "thcgw_eri_read_kpts",
               "thcgw_eri_write")


class EriInfo(ctypes.Structure):
    """thcgw_eri_info"""
    _fields_ = [(k, ctypes.c_int) for k in ("Np", "nbnd", "x_first", "x_last", "y_first", "y_last", "ns_npol",
                                            "nkpts", "nkpts_ibz", "nqpts_ibz", "has_Y", "nkpts_file", "nqpts")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load_eri_library(path=ERI_LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libthcgw_eri.so is not built (%s); it needs the HDF5 C library (make -C coqui_amd/csrc)" % path)
    lib = ctypes.CDLL(path)
    i, l, p, c = ctypes.c_int, ctypes.c_long, ctypes.c_void_p, ctypes.c_char_p
    sig = {
        "thcgw_eri_read_info": [c, ctypes.POINTER(EriInfo), p, i],
        "thcgw_eri_read_X": [c, i, p, l, p, i],
        "thcgw_eri_read_V": [c, p, l, p, i],
        "thcgw_eri_read_kpts": [c, p, l, p, l, p, i],
        "thcgw_eri_write": [c, ctypes.POINTER(EriInfo), p, p, p, p, p, p, p, p, p, i],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.restype = i
        f.argtypes = args
    _lib = lib
    return lib


def _call(fn, *args):
    err = ctypes.create_string_buffer(512)
    rc = fn(*args, err, 512)
    if rc != 0:
        raise IOError("thc.eri.h5: " + err.value.decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def read_thc_eri(path):
    """Returns dict(info=..., X=..., Y=None or array, V=..., kpts=..., qpts=...)."""
    lib = load_eri_library()
    b = os.fsencode(path)
    info = EriInfo()
    _call(lib.thcgw_eri_read_info, b, ctypes.byref(info))
    d = info.as_dict()
    X = np.empty((d["ns_npol"], d["nkpts"], d["Np"], d["x_last"] - d["x_first"]), np.complex128)
    _call(lib.thcgw_eri_read_X, b, 0, _p(X), 2 * X.size)
    Y = None
    if d["has_Y"]:
        Y = np.empty((d["ns_npol"], d["nkpts"], d["Np"], d["y_last"] - d["y_first"]), np.complex128)
        _call(lib.thcgw_eri_read_X, b, 1, _p(Y), 2 * Y.size)
    V = np.empty((d["nqpts_ibz"], d["Np"], d["Np"]), np.complex128)
    _call(lib.thcgw_eri_read_V, b, _p(V), 2 * V.size)
    kpts = np.empty((d["nkpts_file"], 3)) if d["nkpts_file"] else None
    qpts = np.empty((d["nqpts"], 3)) if d["nqpts"] else None
    if kpts is not None or qpts is not None:
        _call(lib.thcgw_eri_read_kpts, b, _p(kpts), 0 if kpts is None else kpts.size, _p(qpts),
              0 if qpts is None else qpts.size)
    return dict(info=d, X=X, Y=Y, V=V, kpts=kpts, qpts=qpts)


def write_thc_eri(path, X, V, kpts=None, qpts=None, x_range=None, nbnd=None, Y=None, y_range=None,
                  interp_points=None, chi_head=None, chi_bar_head=None, nkpts_ibz=None):
    """Write the file in the reference's layout.  X (ns*npol, nk, Np, nx) or (nk, Np, nx) for
    ns*npol = 1, V (nq_ibz, Np, Np)."""
    lib = load_eri_library()
    X = np.ascontiguousarray(X, np.complex128)
    if X.ndim == 3:
        X = X[None]
    V = np.ascontiguousarray(V, np.complex128)
    ns, nk, Np, nx = X.shape
    xr = x_range or (0, nx)
    yr = y_range or xr
    if Y is not None:
        Y = np.ascontiguousarray(Y, np.complex128)
        if Y.ndim == 3:
            Y = Y[None]
    arr = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
    kpts, qpts = arr(kpts, np.float64), arr(qpts, np.float64)
    if kpts is not None and kpts.shape != (nk, 3):
        raise ValueError("kpts: expected shape (%d, 3), got %s" % (nk, kpts.shape))
    if qpts is not None and (qpts.ndim != 2 or qpts.shape[1] != 3):
        raise ValueError("qpts: expected shape (nq, 3), got %s" % (qpts.shape,))
    if Y is not None and Y.shape[:3] != (ns, nk, Np):
        raise ValueError("Y: expected shape (%d, %d, %d, ny), got %s" % (ns, nk, Np, Y.shape))
    if V.ndim != 3 or V.shape[1:] != (Np, Np):
        raise ValueError("V: expected shape (nq_ibz, %d, %d), got %s" % (Np, Np, V.shape))
    info = EriInfo(Np, nbnd if nbnd is not None else -1, xr[0], xr[1], yr[0], yr[1], ns, nk,
                   nkpts_ibz if nkpts_ibz is not None else nk, V.shape[0], int(Y is not None),
                   0 if kpts is None else nk, 0 if qpts is None else qpts.shape[0])
    ip = arr(interp_points, np.int64)
    ch, cb = arr(chi_head, np.complex128), arr(chi_bar_head, np.complex128)
    _call(lib.thcgw_eri_write, os.fsencode(path), ctypes.byref(info), _p(X), _p(Y), _p(V), _p(kpts), _p(qpts), _p(ip),
          _p(ch), _p(cb))
