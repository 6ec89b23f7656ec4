"""Python binding of libthcgw (ctypes over the C ABI in include/thcgw.h).

Device arrays are torch complex128 tensors on the current HIP device; torch provides the
memory and the stream, the library does the work.  The method names follow the reference
routines they replace (scr_coulomb_t::update_w, gw_t::evaluate, gw_t::rpa_energy, ...).

There is no CPU fallback: if libthcgw.so is missing or no GPU is present, construction
raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libthcgw.so")

#: every symbol declared in include/thcgw.h
EXPORTS = (
    "thcgw_create", "thcgw_destroy", "thcgw_last_error", "thcgw_update_G", "thcgw_primary_to_aux",
    "thcgw_primary_to_aux_noalloc",
    "thcgw_aux_to_primary", "thcgw_pi", "thcgw_tau_to_w", "thcgw_dyson_w", "thcgw_w_to_tau", "thcgw_sigma",
    "thcgw_hermitize", "thcgw_rpa_ec", "thcgw_gw_iteration", "thcgw_solve_qp", "thcgw_gm_energy",
    "thcgw_workspace_bytes", "thcgw_set_flag", "thcgw_profile_read", "thcgw_hf", "thcgw_dyson_g",
    "thcgw_nelec", "thcgw_update_mu", "thcgw_batch_create", "thcgw_batch_destroy", "thcgw_batch_last_error",
    "thcgw_batch_set", "thcgw_batch_set_lanes", "thcgw_batch_gw_iteration", "thcgw_pi_kspace", "thcgw_sigma_kspace",
    "thcgw_lri_create", "thcgw_lri_destroy", "thcgw_lri_last_error", "thcgw_lri_set", "thcgw_lri_dyson",
    "thcgw_lri_train", "thcgw_lri_num_pairs", "thcgw_lri_clear_pairs", "thcgw_lri_reset_state",
    "thcgw_mlp_create", "thcgw_mlp_destroy", "thcgw_mlp_last_error", "thcgw_mlp_num_params",
    "thcgw_mlp_get_params", "thcgw_mlp_set_params", "thcgw_mlp_forward", "thcgw_mlp_backward",
    "thcgw_mlp_adam_step", "thcgw_mlp_train_mse", "thcgw_eps_inv_head", "thcgw_sigma_div_correction",
    "thcgw_set_qsym", "thcgw_set_bz", "thcgw_pi_ibz", "thcgw_sigma_ibz", "thcgw_gw_iteration_ibz",
)

ERRORS = {0: "OK", 1: "invalid argument", 2: "HIP error", 3: "singular matrix", 4: "unsupported", 5: "allocation",
          6: "non-finite input"}


class ThcgwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "error"), code, msg))
        self.code = code


class _Dims(ctypes.Structure):
    _fields_ = [("kgrid", ctypes.c_int * 3), ("no", ctypes.c_int), ("nmu", ctypes.c_int), ("ns", ctypes.c_int),
                ("nt", ctypes.c_int), ("nw_b", ctypes.c_int), ("nw_f", ctypes.c_int), ("beta", ctypes.c_double)]


class _Tables(ctypes.Structure):
    _fields_ = [("tau_mesh_f", ctypes.c_void_p), ("wn_mesh_f", ctypes.c_void_p), ("Ttw_ff", ctypes.c_void_p),
                ("Twt_ff", ctypes.c_void_p), ("Ttw_bb", ctypes.c_void_p), ("Twt_bb", ctypes.c_void_p),
                ("T_beta_t_ff", ctypes.c_void_p)]


class _Bz(ctypes.Structure):
    """thcgw_bz"""
    _fields_ = [("nk_ibz", ctypes.c_int), ("nq_ibz", ctypes.c_int), ("nk_trev_pairs", ctypes.c_int),
                ("nsym", ctypes.c_int)] + [(n, ctypes.c_void_p) for n in (
                    "k_grid_index", "kp_to_ibz", "kp_trev", "kp_trev_pair", "qk_to_k2", "qminus", "qp_to_ibz",
                    "qp_trev", "nq_per_s", "Qs", "ks_to_k", "D")]


class LriStats(ctypes.Structure):
    """thcgw_lri_stats"""
    _fields_ = [("solves", ctypes.c_int), ("krylov_accepted", ctypes.c_int), ("fallbacks", ctypes.c_int),
                ("iterations", ctypes.c_int), ("iter_mean", ctypes.c_double), ("resid_max", ctypes.c_double),
                ("core_err", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load_library(path=LIB_PATH):
    """Load libthcgw.so and declare its signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libthcgw.so is not built (%s); run __graft_entry__.build() or make -C coqui_amd/csrc" % path)
    lib = ctypes.CDLL(path)
    i, l, d, p = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
    sig = {
        "thcgw_create": (i, [ctypes.POINTER(_Dims), ctypes.POINTER(_Tables), ctypes.POINTER(p)]),
        "thcgw_destroy": (None, [p]),
        "thcgw_last_error": (ctypes.c_char_p, [p]),
        "thcgw_update_G": (i, [p, p, p, d, p, p]),
        "thcgw_primary_to_aux": (i, [p, l, p, p, p, p]),
        "thcgw_primary_to_aux_noalloc": (i, [p, l, p, p, p, p]),
        "thcgw_aux_to_primary": (i, [p, l, p, p, p, p]),
        "thcgw_pi": (i, [p, p, p, p, p]),
        "thcgw_tau_to_w": (i, [p, p, p, p]),
        "thcgw_pi_kspace": (i, [p, p, p, p, p]),
        "thcgw_sigma_kspace": (i, [p, p, p, p, p, i, p]),
        "thcgw_set_bz": (i, [p, p]),
        "thcgw_pi_ibz": (i, [p, p, p, p, p]),
        "thcgw_sigma_ibz": (i, [p, p, p, p, p, i, p]),
        "thcgw_gw_iteration_ibz": (i, [p, p, p, p, p, p, p, i, p]),
        "thcgw_dyson_w": (i, [p, p, p, p, p, p]),
        "thcgw_w_to_tau": (i, [p, p, p, p]),
        "thcgw_sigma": (i, [p, p, p, p, p, i, p]),
        "thcgw_hermitize": (i, [p, l, i, p, p]),
        "thcgw_rpa_ec": (i, [p, p, p, p, p]),
        "thcgw_gw_iteration": (i, [p, p, p, p, p, p, p, i, p]),
        "thcgw_solve_qp": (i, [p, p, p, p, p, d, i, d, d, p, p]),
        "thcgw_gm_energy": (i, [p, p, p, p, p]),
        "thcgw_hf": (i, [p, p, p, p, i, i, p, p]),
        "thcgw_dyson_g": (i, [p, p, p, p, p, d, p, p, p]),
        "thcgw_nelec": (i, [p, p, p, p, p, d, ctypes.POINTER(d), p]),
        "thcgw_update_mu": (i, [p, p, p, p, p, d, d, d, ctypes.POINTER(d), ctypes.POINTER(d), p]),
        "thcgw_batch_create": (i, [i, ctypes.POINTER(p), ctypes.POINTER(p)]),
        "thcgw_batch_destroy": (None, [p]),
        "thcgw_batch_last_error": (ctypes.c_char_p, [p]),
        "thcgw_batch_set": (i, [p, i, i]),
        "thcgw_batch_set_lanes": (i, [p, i]),
        "thcgw_batch_gw_iteration": (i, [p, ctypes.POINTER(p), ctypes.POINTER(p), ctypes.POINTER(p),
                                         ctypes.POINTER(p), ctypes.POINTER(p), ctypes.POINTER(d), i, p]),
        "thcgw_workspace_bytes": (l, [p]),
        "thcgw_set_flag": (i, [p, i, l]),
        "thcgw_profile_read": (i, [p, ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(l)]),
        "thcgw_eps_inv_head": (i, [p, p, p, p, d, i, p, p, p]),
        "thcgw_set_qsym": (i, [p, p, p, p, p]),
        "thcgw_sigma_div_correction": (i, [p, p, p, p, p, d, p, i, p]),
        "thcgw_lri_create": (i, [p, i, i, ctypes.c_ulonglong, ctypes.POINTER(p)]),
        "thcgw_lri_destroy": (None, [p]),
        "thcgw_lri_last_error": (ctypes.c_char_p, [p]),
        "thcgw_lri_set": (i, [p, i, i, d]),
        "thcgw_lri_dyson": (i, [p, p, p, i, i, ctypes.POINTER(LriStats), p]),
        "thcgw_lri_train": (i, [p, i, i, ctypes.c_float, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float), p]),
        "thcgw_lri_num_pairs": (l, [p]),
        "thcgw_lri_clear_pairs": (i, [p]),
        "thcgw_lri_reset_state": (i, [p]),
        "thcgw_mlp_create": (i, [i, i, i, ctypes.c_ulonglong, ctypes.POINTER(p)]),
        "thcgw_mlp_destroy": (None, [p]),
        "thcgw_mlp_last_error": (ctypes.c_char_p, [p]),
        "thcgw_mlp_num_params": (l, [p]),
        "thcgw_mlp_get_params": (i, [p, p]),
        "thcgw_mlp_set_params": (i, [p, p]),
        "thcgw_mlp_forward": (i, [p, i, p, p, p]),
        "thcgw_mlp_backward": (i, [p, i, p, p, p, p]),
        "thcgw_mlp_adam_step": (i, [p, p, ctypes.c_float, p]),
        "thcgw_mlp_train_mse": (i, [p, i, p, p, ctypes.c_float, ctypes.POINTER(ctypes.c_float), p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(t):
    return ctypes.c_void_p(None if t is None else t.data_ptr())


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _squeezed(shape):
    return tuple(int(d) for d in shape if int(d) != 1)


def check_array(a, shape, name, dtype="complex128", device=True):
    """Raises ValueError unless `a` (torch tensor or numpy array) has exactly `shape` (axes
    of length 1 may be added or dropped), the dtype and C-contiguity, and, for torch
    tensors with device=True, lives on the GPU.  The C ABI takes raw pointers, so this is
    the only place a mis-sized array is caught before the library reads or writes it."""
    shp = tuple(int(d) for d in a.shape)
    if _squeezed(shp) != _squeezed(shape):
        raise ValueError("%s: expected shape %s, got %s" % (name, tuple(int(d) for d in shape), shp))
    dt = str(a.dtype).replace("torch.", "")
    if dt != dtype:
        raise ValueError("%s: expected dtype %s, got %s" % (name, dtype, dt))
    contiguous = a.is_contiguous() if hasattr(a, "is_contiguous") else a.flags["C_CONTIGUOUS"]
    if not contiguous:
        raise ValueError("%s: expected a C-contiguous array" % name)
    if device and hasattr(a, "is_cuda") and not a.is_cuda:
        raise ValueError("%s: expected a GPU tensor, got one on %s" % (name, a.device))
    return a


class ThcGW:
    """One system shape on the current GPU (thcgw_create / thcgw_destroy)."""

    def __init__(self, kgrid, no, nmu, iaft):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("coqui_amd.ThcGW needs a HIP device; there is no CPU fallback")
        self._torch = torch
        self.lib = load_library()
        self.kgrid = tuple(int(v) for v in kgrid)
        self.nk = int(np.prod(self.kgrid))
        self.no, self.nmu = int(no), int(nmu)
        self.iaft = iaft
        dims = _Dims((ctypes.c_int * 3)(*self.kgrid), self.no, self.nmu, 1, iaft.nt, iaft.nw_b, iaft.nw_f,
                     iaft.beta)
        # keep the host tables alive for the duration of the create call
        self._keep = [iaft.tau_mesh, iaft.wn_mesh_f, iaft.Ttw_ff, iaft.Twt_ff, iaft.Ttw_bb, iaft.Twt_bb,
                      iaft.T_beta_t_ff]
        tabs = _Tables(*[_np_ptr(a) for a in self._keep])
        h = ctypes.c_void_p()
        torch.cuda.current_device()  # make sure the HIP runtime is initialised on this device
        rc = self.lib.thcgw_create(ctypes.byref(dims), ctypes.byref(tabs), ctypes.byref(h))
        if rc != 0:
            raise ThcgwError(rc, self.lib.thcgw_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.thcgw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ helpers
    def _check(self, rc):
        if rc != 0:
            raise ThcgwError(rc, self.lib.thcgw_last_error(self.h).decode())

    def _stream(self, stream=None):
        s = stream if stream is not None else self._torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def _dev(self, a, shape, name, dtype="complex128"):
        """`a` as a GPU tensor of `shape`: device tensors are checked (check_array) and used
        in place, host arrays are checked and copied."""
        t = self._torch
        if isinstance(a, t.Tensor):
            return check_array(a, shape, name, dtype)
        h = np.ascontiguousarray(a, np.dtype(dtype))
        check_array(h, shape, name, dtype)
        return t.from_numpy(h).cuda()

    def _out(self, a, shape, name):
        """A caller-given output tensor, checked, or a new one of `shape`."""
        return self._empty(*shape) if a is None else check_array(a, shape, name)

    def _empty(self, *shape):
        return self._torch.empty(shape, dtype=self._torch.complex128, device="cuda")

    # array shapes of the C ABI (include/thcgw.h)
    def _g(self):
        return (self.nt, 1, self.nk, self.no, self.no)

    def _x(self):
        return (1, self.nk, self.nmu, self.no)

    def _v(self):
        return (self.nk, self.nmu, self.nmu)

    def _kij(self):
        return (1, self.nk, self.no, self.no)

    def _pit(self):
        return (self.nt_half, self.nk, self.nmu, self.nmu)

    def _piw(self):
        return (self.nw_half, self.nk, self.nmu, self.nmu)

    def _wqt(self):
        return (self.nk, self.nt_half, self.nmu, self.nmu)

    def _nb(self, O, per, name):
        n = int(np.prod(O.shape))
        if n == 0 or n % per:
            raise ValueError("%s: %d elements is not a multiple of %d" % (name, n, per))
        return n // per

    @property
    def nt(self):
        return self.iaft.nt

    @property
    def nt_half(self):
        return self.iaft.nt_half

    @property
    def nw_half(self):
        return self.iaft.nw_half

    def set_flag(self, flag, value):
        self._check(self.lib.thcgw_set_flag(self.h, int(flag), int(value)))

    def set_qsym(self, src=None, perm=None, phase=None, conj=None):
        """q-point symmetry map (thcgw_set_qsym); src=None clears it."""
        if src is None:
            self._check(self.lib.thcgw_set_qsym(self.h, None, None, None, None))
            return
        self._qsym = [np.ascontiguousarray(src, np.int32), np.ascontiguousarray(perm, np.int32),
                      None if phase is None else np.ascontiguousarray(phase, np.complex128),
                      None if conj is None else np.ascontiguousarray(conj, np.int32)]
        self._check(self.lib.thcgw_set_qsym(self.h, *[None if a is None else _np_ptr(a) for a in self._qsym]))

    def set_bz(self, bz=None):
        """Brillouin-zone symmetry (thcgw_set_bz) from a dict of the tables of
        coqui_amd.system.make_system_bz ("order" is the k_grid_index); None clears it."""
        if bz is None:
            self._check(self.lib.thcgw_set_bz(self.h, None))
            self._bz = None
            return
        ints = {n: np.ascontiguousarray(bz[n], np.int32) for n in (
            "kp_to_ibz", "kp_trev", "kp_trev_pair", "qk_to_k2", "qminus", "qp_to_ibz", "qp_trev", "nq_per_s", "Qs",
            "ks_to_k")}
        ints["k_grid_index"] = np.ascontiguousarray(bz["order"], np.int32)
        D = np.ascontiguousarray(bz["D"], np.complex128)
        nsym = int(len(ints["nq_per_s"]))
        nk_ibz, nq_ibz = int(bz["nk_ibz"]), int(bz["nq_ibz"])
        for name, shape in (("k_grid_index", (self.nk,)), ("kp_to_ibz", (self.nk,)), ("kp_trev", (self.nk,)),
                            ("kp_trev_pair", (self.nk,)), ("qk_to_k2", (self.nk, self.nk)), ("qminus", (self.nk,)),
                            ("qp_to_ibz", (self.nk,)), ("qp_trev", (self.nk,)), ("Qs", (nsym, self.nk)),
                            ("ks_to_k", (nsym, nk_ibz))):
            if ints[name].shape != shape:
                raise ValueError("bz[%s]: expected shape %s, got %s" % (name, shape, ints[name].shape))
        if D.shape != (nsym, nk_ibz, self.no, self.no):
            raise ValueError("bz[D]: expected shape %s, got %s" % ((nsym, nk_ibz, self.no, self.no), D.shape))
        keep = dict(ints, D=D)
        st = _Bz(nk_ibz, nq_ibz, int(bz["nk_trev_pairs"]), nsym,
                 *[_np_ptr(keep[n]) for n in ("k_grid_index", "kp_to_ibz", "kp_trev", "kp_trev_pair", "qk_to_k2",
                                              "qminus", "qp_to_ibz", "qp_trev", "nq_per_s", "Qs", "ks_to_k", "D")])
        self._check(self.lib.thcgw_set_bz(self.h, ctypes.byref(st)))
        self._bz = (nk_ibz, nq_ibz)

    def _bz_dims(self):
        if getattr(self, "_bz", None) is None:
            raise ValueError("no Brillouin-zone symmetry set (set_bz)")
        return self._bz

    def pi_ibz(self, G, X, out=None, stream=None):
        """chi0 on the irreducible q (rpa_pi.icc:22-180 with symmetry): G on the irreducible k."""
        nki, nqi = self._bz_dims()
        G = self._dev(G, (self.nt, 1, nki, self.no, self.no), "G")
        X = self._dev(X, self._x(), "X")
        P = self._out(out, (self.nt_half, nqi, self.nmu, self.nmu), "out")
        self._check(self.lib.thcgw_pi_ibz(self.h, _ptr(G), _ptr(X), _ptr(P), self._stream(stream)))
        return P

    def sigma_ibz(self, G, W_qt, X, hermitize=True, out=None, stream=None):
        """Sigma on the irreducible k (thc_gw.icc:216-443)."""
        nki, nqi = self._bz_dims()
        g = (self.nt, 1, nki, self.no, self.no)
        G, X = self._dev(G, g, "G"), self._dev(X, self._x(), "X")
        W_qt = self._dev(W_qt, (nqi, self.nt_half, self.nmu, self.nmu), "W_qt")
        S = self._out(out, g, "out")
        self._check(self.lib.thcgw_sigma_ibz(self.h, _ptr(G), _ptr(W_qt), _ptr(X), _ptr(S), int(bool(hermitize)),
                                             self._stream(stream)))
        return S

    def gw_iteration_ibz(self, G, X, V, with_ec=True, hermitize=True, Sigma=None, W_qt=None, stream=None):
        """The iteration on the irreducible wedge: returns (Sigma (irreducible k), W_qtPQ
        (irreducible q), Ec or None)."""
        nki, nqi = self._bz_dims()
        g = (self.nt, 1, nki, self.no, self.no)
        G, X = self._dev(G, g, "G"), self._dev(X, self._x(), "X")
        V = self._dev(V, (nqi, self.nmu, self.nmu), "V")
        S = self._out(Sigma, g, "Sigma")
        W = self._out(W_qt, (nqi, self.nt_half, self.nmu, self.nmu), "W_qt")
        ec = ctypes.c_double()
        self._check(self.lib.thcgw_gw_iteration_ibz(self.h, _ptr(G), _ptr(X), _ptr(V), _ptr(S), _ptr(W),
                                                    ctypes.byref(ec) if with_ec else None, int(bool(hermitize)),
                                                    self._stream(stream)))
        return S, W, (ec.value if with_ec else None)

    def profile_read(self):
        """{class: (ms, flops, launches, bytes)} of the GEMM launches recorded since the last
        read (flag 4); class = "NN", "NT", ..., "CC" (op of A then op of B) for FP64 products,
        "NN32", "NC32", "CN32", "CC32" for FP32 products and "NNbf" ... "CCbf" for split-BF16
        products; flops and bytes are the library's algorithmic models."""
        ms = (ctypes.c_double * 17)()
        fl = (ctypes.c_double * 17)()
        by = (ctypes.c_double * 17)()
        n = (ctypes.c_long * 17)()
        self._check(self.lib.thcgw_profile_read(self.h, ms, fl, by, n))
        names = ([a + b for a in "NTC" for b in "NTC"] + ["NN32", "NC32", "CN32", "CC32"]
                 + ["NNbf", "NCbf", "CNbf", "CCbf"])
        return {names[c]: (ms[c], fl[c], n[c], by[c]) for c in range(17) if n[c]}

    def workspace_bytes(self):
        return int(self.lib.thcgw_workspace_bytes(self.h))

    # ------------------------------------------------------------------ stages
    def update_G(self, C, E, mu=0.0, out=None, stream=None):
        C = self._dev(C, (self.nk, self.no, self.no), "C")
        E = self._dev(E, (self.nk, self.no), "E", "float64")
        G = self._out(out, self._g(), "out")
        self._check(self.lib.thcgw_update_G(self.h, _ptr(C), _ptr(E), ctypes.c_double(mu), _ptr(G), self._stream(stream)))
        return G

    def primary_to_aux(self, X, O, stream=None):
        nb = self._nb(O, self.nk * self.no * self.no, "O")
        X, O = self._dev(X, self._x(), "X"), self._dev(O, (nb, self.nk, self.no, self.no), "O")
        out = self._empty(nb, self.nk, self.nmu, self.nmu)
        self._check(self.lib.thcgw_primary_to_aux(self.h, nb, _ptr(X), _ptr(O), _ptr(out), self._stream(stream)))
        return out

    def aux_to_primary(self, X, O, stream=None):
        nb = self._nb(O, self.nk * self.nmu * self.nmu, "O")
        X, O = self._dev(X, self._x(), "X"), self._dev(O, (nb, self.nk, self.nmu, self.nmu), "O")
        out = self._empty(nb, self.nk, self.no, self.no)
        self._check(self.lib.thcgw_aux_to_primary(self.h, nb, _ptr(X), _ptr(O), _ptr(out), self._stream(stream)))
        return out

    def pi(self, G, X, out=None, stream=None):
        G, X = self._dev(G, self._g(), "G"), self._dev(X, self._x(), "X")
        P = self._out(out, self._pit(), "out")
        self._check(self.lib.thcgw_pi(self.h, _ptr(G), _ptr(X), _ptr(P), self._stream(stream)))
        return P

    def pi_kspace(self, G, X, out=None, stream=None):
        """k-space chi0 (rpa_pi.icc:182-336): Pi^k_PQ(q) = Pi^R_QP(-q) for Hermitian G."""
        G, X = self._dev(G, self._g(), "G"), self._dev(X, self._x(), "X")
        P = self._out(out, self._pit(), "out")
        self._check(self.lib.thcgw_pi_kspace(self.h, _ptr(G), _ptr(X), _ptr(P), self._stream(stream)))
        return P

    def sigma_kspace(self, G, W_qt, X, hermitize=True, out=None, stream=None):
        """k-space self-energy (thc_gw.icc:342-443)."""
        G, W_qt, X = self._dev(G, self._g(), "G"), self._dev(W_qt, self._wqt(), "W_qt"), self._dev(X, self._x(), "X")
        S = self._out(out, self._g(), "out")
        self._check(self.lib.thcgw_sigma_kspace(self.h, _ptr(G), _ptr(W_qt), _ptr(X), _ptr(S), int(bool(hermitize)),
                                                self._stream(stream)))
        return S

    def tau_to_w(self, Pi_t, out=None, stream=None):
        Pi_t = self._dev(Pi_t, self._pit(), "Pi_t")
        P = self._out(out, self._piw(), "out")
        self._check(self.lib.thcgw_tau_to_w(self.h, _ptr(Pi_t), _ptr(P), self._stream(stream)))
        return P

    def dyson_w(self, PiW, V, with_terms=False, stream=None):
        """In place: PiW <- W - V.  Returns (trace, logdet) host arrays when with_terms."""
        PiW, V = self._dev(PiW, self._piw(), "PiW"), self._dev(V, self._v(), "V")
        n = self.nw_half * self.nk
        tr = np.zeros(n, np.complex128) if with_terms else None
        ld = np.zeros(n, np.float64) if with_terms else None
        self._check(self.lib.thcgw_dyson_w(self.h, _ptr(PiW), _ptr(V), _np_ptr(tr) if with_terms else None,
                                           _np_ptr(ld) if with_terms else None, self._stream(stream)))
        return (tr, ld) if with_terms else PiW

    def w_to_tau(self, W_w, out=None, stream=None):
        W_w = self._dev(W_w, self._piw(), "W_w")
        W = self._out(out, self._wqt(), "out")
        self._check(self.lib.thcgw_w_to_tau(self.h, _ptr(W_w), _ptr(W), self._stream(stream)))
        return W

    def sigma(self, G, W_qt, X, hermitize=True, out=None, stream=None):
        G, W_qt, X = self._dev(G, self._g(), "G"), self._dev(W_qt, self._wqt(), "W_qt"), self._dev(X, self._x(), "X")
        S = self._out(out, self._g(), "out")
        self._check(self.lib.thcgw_sigma(self.h, _ptr(G), _ptr(W_qt), _ptr(X), _ptr(S), int(bool(hermitize)),
                                         self._stream(stream)))
        return S

    def hermitize(self, A, stream=None):
        if len(A.shape) < 2 or A.shape[-1] != A.shape[-2]:
            raise ValueError("A: expected a stack of square matrices, got shape %s" % (tuple(A.shape),))
        n = int(A.shape[-1])
        A = self._dev(A, (self._nb(A, n * n, "A"), n, n), "A")
        self._check(self.lib.thcgw_hermitize(self.h, A.numel() // (n * n), n, _ptr(A), self._stream(stream)))
        return A

    def rpa_ec(self, Pi_w, V, stream=None):
        Pi_w, V = self._dev(Pi_w, self._piw(), "Pi_w"), self._dev(V, self._v(), "V")
        ec = ctypes.c_double()
        self._check(self.lib.thcgw_rpa_ec(self.h, _ptr(Pi_w), _ptr(V), ctypes.byref(ec), self._stream(stream)))
        return ec.value

    def gw_iteration(self, G, X, V, with_ec=True, hermitize=True, Sigma=None, W_qt=None, stream=None):
        """update_w + evaluate: returns (Sigma_tskij, W_qtPQ, Ec or None)."""
        G, X, V = self._dev(G, self._g(), "G"), self._dev(X, self._x(), "X"), self._dev(V, self._v(), "V")
        S = self._out(Sigma, self._g(), "Sigma")
        W = self._out(W_qt, self._wqt(), "W_qt")
        ec = ctypes.c_double()
        self._check(self.lib.thcgw_gw_iteration(self.h, _ptr(G), _ptr(X), _ptr(V), _ptr(S), _ptr(W),
                                                ctypes.byref(ec) if with_ec else None, int(bool(hermitize)),
                                                self._stream(stream)))
        return S, W, (ec.value if with_ec else None)

    def solve_qp(self, Sigma, C, Vhf, E0, mu=0.0, nfit=18, eta=1e-4, tol=1e-8, stream=None):
        Sigma, C = self._dev(Sigma, self._g(), "Sigma"), self._dev(C, (self.nk, self.no, self.no), "C")
        Vhf = self._dev(Vhf, (self.nk, self.no, self.no), "Vhf")
        E0 = check_array(np.ascontiguousarray(E0, np.float64), (self.nk, self.no), "E0", "float64")
        E = np.zeros_like(E0)
        self._check(self.lib.thcgw_solve_qp(self.h, _ptr(Sigma), _ptr(C), _ptr(Vhf), _np_ptr(E0), ctypes.c_double(mu),
                                            int(nfit), ctypes.c_double(eta), ctypes.c_double(tol), _np_ptr(E),
                                            self._stream(stream)))
        return E

    def hf(self, Dm, X, V, hartree=True, exchange=True, out=None, stream=None):
        """THC Hartree-Fock potential F_skij (1, nk, no, no) of the density matrix Dm_skij."""
        Dm, X, V = self._dev(Dm, self._kij(), "Dm"), self._dev(X, self._x(), "X"), self._dev(V, self._v(), "V")
        F = self._out(out, self._kij(), "out")
        self._check(self.lib.thcgw_hf(self.h, _ptr(Dm), _ptr(X), _ptr(V), int(bool(hartree)), int(bool(exchange)),
                                      _ptr(F), self._stream(stream)))
        return F

    def _opt_dev(self, a, shape, name):
        return None if a is None else self._dev(a, shape, name)

    def _dyson_g_args(self, F, Sigma, H0, S):
        return (self._dev(F, self._kij(), "F"), self._dev(Sigma, self._g(), "Sigma"),
                self._opt_dev(H0, self._kij(), "H0"), self._opt_dev(S, self._kij(), "S"))

    def dyson_g(self, F, Sigma, mu, H0=None, S=None, want_dm=True, out=None, stream=None):
        """Matsubara Dyson equation for G: returns (G_tskij, Dm_skij or None)."""
        F, Sigma, H0, S = self._dyson_g_args(F, Sigma, H0, S)
        G = self._out(out, self._g(), "out")
        Dm = self._empty(1, self.nk, self.no, self.no) if want_dm else None
        self._check(self.lib.thcgw_dyson_g(self.h, _ptr(H0), _ptr(F), _ptr(S), _ptr(Sigma), ctypes.c_double(mu),
                                           _ptr(G), _ptr(Dm), self._stream(stream)))
        return G, Dm

    def nelec(self, F, Sigma, mu, H0=None, S=None, stream=None):
        F, Sigma, H0, S = self._dyson_g_args(F, Sigma, H0, S)
        n = ctypes.c_double()
        self._check(self.lib.thcgw_nelec(self.h, _ptr(H0), _ptr(F), _ptr(S), _ptr(Sigma), ctypes.c_double(mu),
                                         ctypes.byref(n), self._stream(stream)))
        return n.value

    def update_mu(self, F, Sigma, nel_target, old_mu, mu_tol=1e-9, H0=None, S=None, stream=None):
        """Chemical potential for nel_target: returns (mu, nelec)."""
        F, Sigma, H0, S = self._dyson_g_args(F, Sigma, H0, S)
        mu, n = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.thcgw_update_mu(self.h, _ptr(H0), _ptr(F), _ptr(S), _ptr(Sigma),
                                             ctypes.c_double(nel_target), ctypes.c_double(old_mu),
                                             ctypes.c_double(mu_tol), ctypes.byref(mu), ctypes.byref(n),
                                             self._stream(stream)))
        return mu.value, n.value

    DIV = {"ignore_g0": 0, "gygi": 1, "gygi_extrplt": 2, "gygi_extrplt_2d": 3}

    def eps_inv_head(self, W_qt, chi_bar, qpts_cart, volume, div_treatment="gygi_extrplt", stream=None):
        """Head of the inverse dielectric function: returns (eps_inv_tq (nt_half, nk), eps_inv_t (nt_half))."""
        W_qt, chi_bar = self._dev(W_qt, self._wqt(), "W_qt"), self._dev(chi_bar, (self.nk, self.nmu), "chi_bar")
        q = check_array(np.ascontiguousarray(qpts_cart, np.float64), (self.nk, 3), "qpts_cart", "float64")
        tq = np.zeros((self.nt_half, self.nk), np.complex128)
        t0 = np.zeros(self.nt_half, np.complex128)
        mode = self.DIV[div_treatment] if isinstance(div_treatment, str) else int(div_treatment)
        self._check(self.lib.thcgw_eps_inv_head(self.h, _ptr(W_qt), _ptr(chi_bar), _np_ptr(q), ctypes.c_double(volume),
                                                mode, _np_ptr(tq), _np_ptr(t0), self._stream(stream)))
        return tq, t0

    def sigma_div_correction(self, Sigma, G, X, chi_head, madelung, eps_inv_t, div_treatment="gygi_extrplt",
                             stream=None):
        """In place: Sigma += the long-wavelength correction (thc_gw.icc:446-526)."""
        Sigma, G = self._dev(Sigma, self._g(), "Sigma"), self._dev(G, self._g(), "G")
        X, chi_head = self._dev(X, self._x(), "X"), self._dev(chi_head, (self.nmu,), "chi_head")
        e = check_array(np.ascontiguousarray(eps_inv_t, np.complex128), (self.nt_half,), "eps_inv_t")
        mode = self.DIV[div_treatment] if isinstance(div_treatment, str) else int(div_treatment)
        self._check(self.lib.thcgw_sigma_div_correction(self.h, _ptr(Sigma), _ptr(G), _ptr(X), _ptr(chi_head),
                                                        ctypes.c_double(madelung), _np_ptr(e), mode,
                                                        self._stream(stream)))
        return Sigma

    def gm_energy(self, G, Sigma, stream=None):
        G, Sigma = self._dev(G, self._g(), "G"), self._dev(Sigma, self._g(), "Sigma")
        e = ctypes.c_double()
        self._check(self.lib.thcgw_gm_energy(self.h, _ptr(G), _ptr(Sigma), ctypes.byref(e), self._stream(stream)))
        return e.value


class ThcBatch:
    """Independent systems (one ThcGW each, uneven sizes) iterated together: thcgw_batch_*.

    The contexts share one scratch pool while the batch lives (one per lane with lanes > 1);
    close the batch before its contexts.  The first iteration runs eagerly (and tunes the
    FP32 tile shape of the mixed-precision systems); later ones replay one captured HIP
    graph, whose lanes run as parallel branches."""

    def __init__(self, contexts, use_graph=True, tune=True, lanes=1):
        self.ctxs = list(contexts)
        self.lib = load_library()
        self._torch = self.ctxs[0]._torch
        arr = (ctypes.c_void_p * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = ctypes.c_void_p()
        rc = self.lib.thcgw_batch_create(len(self.ctxs), arr, ctypes.byref(h))
        if rc != 0:
            raise ThcgwError(rc, self.lib.thcgw_batch_last_error(None).decode())
        self.h = h
        self._check(self.lib.thcgw_batch_set(self.h, int(bool(use_graph)), int(bool(tune))))
        if lanes != 1:
            self.set_lanes(lanes)

    def set_lanes(self, lanes):
        """Spread the systems over `lanes` streams (parallel branches of the graph), before
        the first iteration: thcgw_batch_set_lanes."""
        self._check(self.lib.thcgw_batch_set_lanes(self.h, int(lanes)))

    def _check(self, rc):
        if rc != 0:
            raise ThcgwError(rc, self.lib.thcgw_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.thcgw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gw_iteration(self, G, X, V, Sigma, W_qt, with_ec=True, hermitize=True, stream=None):
        """One iteration of every system; G..W_qt are lists of device tensors (one per system).
        Returns the list of Ec values (or None)."""
        n = len(self.ctxs)
        for lst in (G, X, V, Sigma, W_qt):
            if len(lst) != n:
                raise ValueError("one array per system: %d systems, %d arrays" % (n, len(lst)))
        for i, (c, g, x, v, sg, w) in enumerate(zip(self.ctxs, G, X, V, Sigma, W_qt)):
            for t, shape, name in ((g, c._g(), "G"), (x, c._x(), "X"), (v, c._v(), "V"), (sg, c._g(), "Sigma"),
                                   (w, c._wqt(), "W_qt")):
                if not isinstance(t, self._torch.Tensor):
                    raise ValueError("system %d %s: expected a GPU tensor" % (i, name))
                check_array(t, shape, "system %d %s" % (i, name))
        mk = lambda lst: (ctypes.c_void_p * n)(*[t.data_ptr() for t in lst])
        ec = (ctypes.c_double * n)() if with_ec else None
        s = stream if stream is not None else self._torch.cuda.current_stream()
        self._check(self.lib.thcgw_batch_gw_iteration(self.h, mk(G), mk(X), mk(V), mk(Sigma), mk(W_qt), ec,
                                                       int(bool(hermitize)), ctypes.c_void_p(s.cuda_stream)))
        return list(ec) if with_ec else None


class LearnedDyson:
    """Learned low-rank approximate inverse for the Dyson solve (thcgw_lri_*): a Krylov
    (preconditioned CG, FP64) replacement of the direct Hermitian solve, with the exact
    direct solve as fallback whenever a column's true residual exceeds tol.

    mode 0: no preconditioner; 1: Woodbury core on the sketch basis; 2: previous exact
    core; 3: previous core plus the network's correction (the learned form)."""

    MODES = {"none": 0, "woodbury": 1, "previous": 2, "learned": 3}

    def __init__(self, ctx, rank=64, width=512, seed=20261202):
        self.ctx = ctx
        self.lib = ctx.lib
        self._torch = ctx._torch
        self.rank, self.width = int(rank), int(width)
        h = ctypes.c_void_p()
        rc = self.lib.thcgw_lri_create(ctx.h, self.rank, self.width, ctypes.c_ulonglong(seed), ctypes.byref(h))
        if rc != 0:
            raise ThcgwError(rc, self.lib.thcgw_last_error(ctx.h).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise ThcgwError(rc, self.lib.thcgw_lri_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.thcgw_lri_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, mode="learned", max_iter=200, tol=1e-10):
        m = self.MODES[mode] if isinstance(mode, str) else int(mode)
        self._check(self.lib.thcgw_lri_set(self.h, m, int(max_iter), ctypes.c_double(tol)))

    def dyson(self, PiW, V, krylov=True, record=False, keep_state=False, stream=None):
        """In place PiW <- W - V; returns the call's statistics (dict).  record stores
        training pairs; keep_state leaves the previous-iteration state untouched."""
        PiW, V = self.ctx._dev(PiW, self.ctx._piw(), "PiW"), self.ctx._dev(V, self.ctx._v(), "V")
        st = LriStats()
        flags = int(bool(record)) | (2 if keep_state else 0)
        self._check(self.lib.thcgw_lri_dyson(self.h, _ptr(PiW), _ptr(V), int(bool(krylov)), flags,
                                             ctypes.byref(st), self.ctx._stream(stream)))
        return st.as_dict()

    def train(self, epochs=30, batch=256, lr=1e-3, seed=7, stream=None):
        """Adam on the recorded pairs; returns (mean loss of the first epoch, of the last)."""
        f, l = ctypes.c_float(), ctypes.c_float()
        self._check(self.lib.thcgw_lri_train(self.h, int(epochs), int(batch), ctypes.c_float(lr),
                                             ctypes.c_ulonglong(seed), ctypes.byref(f), ctypes.byref(l),
                                             self.ctx._stream(stream)))
        return f.value, l.value

    def num_pairs(self):
        return int(self.lib.thcgw_lri_num_pairs(self.h))

    def clear_pairs(self):
        self._check(self.lib.thcgw_lri_clear_pairs(self.h))

    def reset_state(self):
        self._check(self.lib.thcgw_lri_reset_state(self.h))


class Mlp:
    """The network of the learned form on its own (thcgw_mlp_*), for checks of its kernels.
    Inputs and outputs are float32 torch tensors on the GPU."""

    def __init__(self, n_in, width, n_out, seed=1):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("coqui_amd.Mlp needs a HIP device; there is no CPU fallback")
        self._torch = torch
        self.lib = load_library()
        torch.cuda.current_device()
        h = ctypes.c_void_p()
        rc = self.lib.thcgw_mlp_create(int(n_in), int(width), int(n_out), ctypes.c_ulonglong(seed), ctypes.byref(h))
        if rc != 0:
            raise ThcgwError(rc, self.lib.thcgw_mlp_last_error(None).decode())
        self.h = h
        self.n_in, self.width, self.n_out = int(n_in), int(width), int(n_out)
        self.n_params = int(self.lib.thcgw_mlp_num_params(h))

    def _check(self, rc):
        if rc != 0:
            raise ThcgwError(rc, self.lib.thcgw_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.thcgw_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        return ctypes.c_void_p(self._torch.cuda.current_stream().cuda_stream)

    def params(self):
        """FP32 master parameters W0, b0, W1, b1, W2, b2 as a flat numpy array."""
        h = np.empty(self.n_params, np.float32)
        self._check(self.lib.thcgw_mlp_get_params(self.h, _np_ptr(h)))
        return h

    def set_params(self, flat):
        flat = check_array(np.ascontiguousarray(flat, np.float32), (self.n_params,), "flat", "float32")
        self._check(self.lib.thcgw_mlp_set_params(self.h, _np_ptr(flat)))

    def _f32(self, a, shape, name):
        t = self._torch
        if isinstance(a, t.Tensor):
            return check_array(a, shape, name, "float32")
        h = check_array(np.ascontiguousarray(a, np.float32), shape, name, "float32")
        return t.from_numpy(h).cuda()

    def _rows(self, X):
        if len(X.shape) != 2:
            raise ValueError("X: expected (batch, %d), got shape %s" % (self.n_in, tuple(X.shape)))
        return int(X.shape[0])

    def forward(self, X):
        X = self._f32(X, (self._rows(X), self.n_in), "X")
        Y = self._torch.empty((X.shape[0], self.n_out), dtype=self._torch.float32, device="cuda")
        self._check(self.lib.thcgw_mlp_forward(self.h, int(X.shape[0]), _ptr(X), _ptr(Y), self._stream()))
        return Y

    def backward(self, X, dY):
        """Gradient of sum(Y * dY) w.r.t. the flat parameters, after forward(X)."""
        b = self._rows(X)
        X, dY = self._f32(X, (b, self.n_in), "X"), self._f32(dY, (b, self.n_out), "dY")
        g = self._torch.empty(self.n_params, dtype=self._torch.float32, device="cuda")
        self._check(self.lib.thcgw_mlp_backward(self.h, int(X.shape[0]), _ptr(X), _ptr(dY), _ptr(g), self._stream()))
        return g

    def adam_step(self, grad, lr):
        grad = self._f32(grad, (self.n_params,), "grad")
        self._check(self.lib.thcgw_mlp_adam_step(self.h, _ptr(grad), ctypes.c_float(lr), self._stream()))

    def train_mse(self, X, T, lr):
        b = self._rows(X)
        X, T = self._f32(X, (b, self.n_in), "X"), self._f32(T, (b, self.n_out), "T")
        loss = ctypes.c_float()
        self._check(self.lib.thcgw_mlp_train_mse(self.h, int(X.shape[0]), _ptr(X), _ptr(T), ctypes.c_float(lr),
                                                 ctypes.byref(loss), self._stream()))
        return loss.value
