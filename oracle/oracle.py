"""FP64 CPU oracle for the THC GW/RPA inner loop -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker (or, in bench.py, as the timed CPU baseline).  The
product (coqui_amd/, include/) never loads it.

The arithmetic lives in thc_oracle.cpp (plain C++/OpenMP, its own ZGEMM and partial-
pivot LU, built by oracle/Makefile into oracle/_build/liboracle.so).  This module wraps
it for numpy arrays and restates the IR-table handling of the reference.

Parity pins (what checks each oracle function; see tests/test_oracle_pins.py):

  kR_coefficients (kpoint_utils.hpp:461-545,588-680)   f_kR f_Rk = I; f_Rk x equals
                                                        numpy.fft.fftn(x)/N_k on the k-grid
  update_G        (qp_scf_common.cpp:82-118)           closed form per orbital (single pole):
                                                        tau_to_w gives 1/(iw_n - e); tau_to_beta
                                                        gives -C f C^H (Fermi occupations)
  primary_to_aux / aux_to_primary (thc_solver_comm.hpp) brute-force 4-index einsum; X = [I;0]
  pi_rspace       (rpa_pi.icc:22-180)                  the reference's k-space algorithm
                                                        (pi_kspace, rpa_pi.icc:182-336) via
                                                        Pi^k_PQ(q) = Pi^R_QP(-q); N_k = 1 and
                                                        N_mu = 1 brute force; Hermiticity in PQ
  tau_to_w_phsym / w_to_tau_phsym (IAFT.icc:49-128)    the reference's IAFT golden data
                                                        (test_iaft.cpp:46-119 fixture) through
                                                        the full-grid transforms; PH folding vs
                                                        the unfolded transform of a PH-symmetric
                                                        function; Ttw Twt = I (test_iaft.cpp:40-43)
  dyson           (scr_coulomb_t.cpp:121-206)          N_mu = 1 closed form V/(1 - V Pi) - V;
                                                        numpy.linalg.inv on small shapes; residual
  sigma_rspace    (thc_gw.icc:70-213)                  sigma_kspace (thc_gw.icc:342-443);
                                                        brute-force convolution on tiny shapes
  rpa_ec          (thc_rpa.icc:72-181)                 N_mu = 1 closed form through the IR sums;
                                                        numpy slogdet
  pade_eval       (pade_t.hpp:34-168)                  exact reproduction of a rational function
  solve_qp        (qp_scf_common.cpp:150-311,679-728)  known root of a single-pole self-energy
  gm_energy       (scf_common.cpp:87-136)              closed form for diagonal G and Sigma
  dyson_g         (simple_dyson.cpp:11-127)            Sigma = 0, F = H(k): G equals update_G's
                                                        closed form and Dm = C f C^H
  nelec / update_mu (scf_common.cpp:10-42,147-202)     numpy eigenvalues of H + Sigma(iw_n) (the
                                                        reference's compute_eigenspectra);
                                                        Sigma = 0: 2 sum_k w_k sum_b f(e_b - mu)
                                                        and the mu that solves it
  eps_inv_head    (g0_div_utils.hpp:55-111,180-322)    real for Hermitian W (a misplaced conjugate is
                                                        not); W = a I closed form; the q -> 0 fit
                                                        reproduces a quadratic in |q| exactly; the
                                                        closest-q and few-q rules
  sigma_div_correction (thc_gw.icc:446-526)            X = [I; 0] closed form conj(h_i) G_ij h_j;
                                                        tau and beta - tau share the head
  thc_hf          (thc_hf.icc:18-300)                  N_k = 1 closed form X^H[diag(J) - D^P o V]X;
                                                        brute-force k - q convolution of the
                                                        exchange; Hermiticity; negative exchange
                                                        energy

No function of this oracle is left "parity unpinned"; the reference's own end-to-end
values (test_thc_gw.cpp) need the ISDF/QE/PySCF stack and cannot be used as pins.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        d = ctypes.c_double
        i = ctypes.c_int
        l = ctypes.c_long
        p = ctypes.c_void_p
        _lib.orc_rpa_ec.restype = d
        _lib.orc_gm_energy.restype = d
        _lib.orc_dyson.restype = i
        _lib.orc_dyson_g.restype = i
        _lib.orc_nelec.restype = d
        _lib.orc_update_mu.restype = d
        _lib.orc_zinverse.restype = i
        _lib.orc_num_threads.restype = i
        for f in ("orc_update_G", "orc_primary_to_aux", "orc_aux_to_primary", "orc_pi_rspace", "orc_pi_kspace",
                  "orc_tau_to_w_phsym", "orc_w_to_tau_phsym", "orc_transform", "orc_sigma_rspace",
                  "orc_sigma_kspace", "orc_thc_hf", "orc_hermitize", "orc_pade_eval", "orc_solve_qp", "orc_kR_coefficients",
                  "orc_zgemm", "orc_set_num_threads"):
            getattr(_lib, f).restype = None
        del d, i, l, p
    return _lib


def _c(a):
    """Contiguous complex128 array (no copy when already so)."""
    return np.ascontiguousarray(a, dtype=np.complex128)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    lib().orc_set_num_threads(ctypes.c_int(int(n)))


def num_threads():
    return int(lib().orc_num_threads())


# ---------------------------------------------------------------- IR tables
def load_ir(beta, nt, prec="high", table=None):
    """IR tables of the reference, rescaled as ir_driver.hpp:68-73.

    The table is chosen by name (e.g. "1e2.1e-15") or by the number of tau points.
    """
    d = os.path.join(os.path.dirname(_HERE), "coqui_amd", "data", "ir")
    if table is None:
        for fn in sorted(os.listdir(d)):
            if fn.endswith(".npz"):
                z = np.load(os.path.join(d, fn))
                if int(z["f_nt"]) == nt and int(z["b_nt"]) == nt:
                    table = fn[:-4]
                    break
    z = np.load(os.path.join(d, table + ".npz"))
    s2 = np.sqrt(2.0)
    t = {
        "name": table,
        "beta": float(beta),
        "nt": int(z["f_nt"]),
        "nw_f": int(z["f_nw"]),
        "nw_b": int(z["b_nw"]),
        "x": z["f_tau_mesh"].copy(),
        "wn_f": z["f_wn_mesh"].astype(np.int64),
        "wn_b": z["b_wn_mesh"].astype(np.int64),
        "Ttw_ff": z["f_Ttw"] * (s2 / beta),
        "Twt_ff": z["f_Twt"] * (beta / s2),
        "Ttw_bb": z["b_Ttw"] * (s2 / beta),
        "Twt_bb": z["b_Twt"] * (beta / s2),
        "T_beta": z["f_T_betat"].copy(),
        "T_zero": z["f_T_zerot"].copy(),
        "raw": z,
    }
    t["tau"] = (t["x"] + 1.0) * beta / 2.0
    return t


# ---------------------------------------------------------------- wrappers
def kR_coefficients(kgrid):
    kg = (ctypes.c_int * 3)(*kgrid)
    nk = int(np.prod(kgrid))
    fRk = np.zeros((nk, nk), np.complex128)
    fkR = np.zeros((nk, nk), np.complex128)
    lib().orc_kR_coefficients(kg, _p(fRk), _p(fkR))
    return fRk, fkR


def zgemm(ta, tb, A, B):
    """op(A) @ op(B) through the oracle's own ZGEMM (for self-checks)."""
    A, B = _c(A), _c(B)
    M = A.shape[1] if ta != "N" else A.shape[0]
    K = A.shape[0] if ta != "N" else A.shape[1]
    N = B.shape[0] if tb != "N" else B.shape[1]
    C = np.zeros((M, N), np.complex128)
    one = np.array([1.0 + 0j])
    zero = np.array([0.0 + 0j])
    lib().orc_zgemm(ctypes.c_char(ta.encode()), ctypes.c_char(tb.encode()), ctypes.c_long(M), ctypes.c_long(N),
                    ctypes.c_long(K), _p(one), _p(A), ctypes.c_long(A.shape[1]), _p(B), ctypes.c_long(B.shape[1]),
                    _p(zero), _p(C), ctypes.c_long(N))
    return C


def update_G(tab, C, E, mu=0.0):
    C = _c(C)
    E = np.ascontiguousarray(E, np.float64)
    nk, no = E.shape
    nt = tab["nt"]
    G = np.zeros((nt, 1, nk, no, no), np.complex128)
    tau = np.ascontiguousarray(tab["tau"], np.float64)
    lib().orc_update_G(nt, nk, no, _p(tau), ctypes.c_double(tab["beta"]), ctypes.c_double(mu), _p(C), _p(E), _p(G))
    return G


def primary_to_aux(X, O):
    X, O = _c(X), _c(O)
    nb, nk, no = O.shape[0], O.shape[1], O.shape[2]
    nmu = X.shape[-2]
    out = np.zeros((nb, nk, nmu, nmu), np.complex128)
    lib().orc_primary_to_aux(ctypes.c_long(nb), nk, nmu, no, _p(X), _p(O), _p(out))
    return out


def aux_to_primary(X, O):
    X, O = _c(X), _c(O)
    nb, nk, nmu = O.shape[0], O.shape[1], O.shape[2]
    no = X.shape[-1]
    out = np.zeros((nb, nk, no, no), np.complex128)
    lib().orc_aux_to_primary(ctypes.c_long(nb), nk, nmu, no, _p(X), _p(O), _p(out))
    return out


def _kg(kgrid):
    return (ctypes.c_int * 3)(*[int(v) for v in kgrid])


def pi_rspace(G, X, kgrid):
    G, X = _c(G), _c(X)
    nt, no = G.shape[0], G.shape[-1]
    nk = int(np.prod(kgrid))
    nmu = X.shape[-2]
    nth = (nt + 1) // 2
    Pi = np.zeros((nth, nk, nmu, nmu), np.complex128)
    lib().orc_pi_rspace(nt, _kg(kgrid), nmu, no, _p(G), _p(X), _p(Pi))
    return Pi


def pi_kspace(G, X, kgrid):
    G, X = _c(G), _c(X)
    nt, no = G.shape[0], G.shape[-1]
    nk = int(np.prod(kgrid))
    nmu = X.shape[-2]
    nth = (nt + 1) // 2
    Pi = np.zeros((nth, nk, nmu, nmu), np.complex128)
    lib().orc_pi_kspace(nt, _kg(kgrid), nmu, no, _p(G), _p(X), _p(Pi))
    return Pi


def tau_to_w_phsym(tab, Xt):
    Xt = _c(Xt)
    nt_b, nw_b = tab["nt"], tab["nw_b"]
    nwh = (nw_b + 1) // 2
    dim = int(np.prod(Xt.shape[1:]))
    Xw = np.zeros((nwh,) + Xt.shape[1:], np.complex128)
    T = _c(tab["Twt_bb"])
    lib().orc_tau_to_w_phsym(nt_b, nw_b, _p(T), ctypes.c_long(dim), _p(Xt), _p(Xw))
    return Xw


def w_to_tau_phsym(tab, Xw):
    Xw = _c(Xw)
    nt_b, nw_b = tab["nt"], tab["nw_b"]
    nth = (nt_b + 1) // 2
    dim = int(np.prod(Xw.shape[1:]))
    Xt = np.zeros((nth,) + Xw.shape[1:], np.complex128)
    T = _c(tab["Ttw_bb"])
    lib().orc_w_to_tau_phsym(nt_b, nw_b, _p(T), ctypes.c_long(dim), _p(Xw), _p(Xt))
    return Xt


def transform(T, X):
    """Full-grid transform out = T @ X over the leading axis (IAFT.icc:9-26, 81-98)."""
    T, X = _c(T), _c(X)
    dim = int(np.prod(X.shape[1:]))
    out = np.zeros((T.shape[0],) + X.shape[1:], np.complex128)
    lib().orc_transform(T.shape[0], T.shape[1], _p(T), ctypes.c_long(dim), _p(X), _p(out))
    return out


def tau_to_beta(tab, Xt):
    """A(beta^-) = T_beta_t_ff . A(tau) (IAFT.icc:226-238)."""
    return np.tensordot(tab["T_beta"], _c(Xt), axes=(0, 0))


def dyson(Piw, V):
    """Returns W - V for Pi(iW) (nw_half, nq, nmu, nmu) and V (nq, nmu, nmu)."""
    W = _c(Piw).copy()
    V = _c(V)
    nwh, nq, nmu = W.shape[0], W.shape[1], W.shape[2]
    info = lib().orc_dyson(nwh, nq, nmu, _p(W), _p(V))
    if info != 0:
        raise np.linalg.LinAlgError("oracle dyson: singular (I - V Pi), info=%d" % info)
    return W


def w_to_qt(Wt):
    """(nt_half, nq, ...) -> (nq, nt_half, ...): the transpose of update_w (scr_coulomb_t.cpp:73-86)."""
    return np.ascontiguousarray(np.swapaxes(Wt, 0, 1))


def sigma_rspace(G, X, Wqt, kgrid):
    G, X, Wqt = _c(G), _c(X), _c(Wqt)
    nt, no = G.shape[0], G.shape[-1]
    nmu = X.shape[-2]
    S = np.zeros_like(G)
    lib().orc_sigma_rspace(nt, _kg(kgrid), nmu, no, _p(G), _p(X), _p(Wqt), _p(S))
    return S


def sigma_kspace(G, X, Wqt, kgrid):
    G, X, Wqt = _c(G), _c(X), _c(Wqt)
    nt, no = G.shape[0], G.shape[-1]
    nmu = X.shape[-2]
    S = np.zeros_like(G)
    lib().orc_sigma_kspace(nt, _kg(kgrid), nmu, no, _p(G), _p(X), _p(Wqt), _p(S))
    return S


def thc_hf(Dm, X, V, kgrid, hartree=True, exchange=True):
    """THC Hartree-Fock potential F_kij for a density matrix Dm (nk, no, no) (thc_hf.icc:18-300)."""
    Dm, X, V = _c(Dm), _c(X), _c(V)
    nk, no = Dm.shape[-3], Dm.shape[-1]
    nmu = X.shape[-2]
    F = np.zeros((nk, no, no), np.complex128)
    lib().orc_thc_hf(_kg(kgrid), nmu, no, _p(Dm), _p(X), _p(V), int(bool(hartree)), int(bool(exchange)), _p(F))
    return F


def _opt(a):
    return None if a is None else _p(_c(a))


def dyson_g(tab, F, Sigma, mu, H0=None, S=None):
    """G(tau) (nt, 1, nk, no, no) and Dm (nk, no, no) from the Matsubara Dyson equation
    (simple_dyson.cpp:11-127)."""
    F, Sigma = _c(F), _c(Sigma)
    H0 = None if H0 is None else _c(H0)
    S = None if S is None else _c(S)
    nt, nk, no = tab["nt"], F.shape[-3], F.shape[-1]
    G = np.zeros((nt, 1, nk, no, no), np.complex128)
    Dm = np.zeros((nk, no, no), np.complex128)
    wn = np.ascontiguousarray(tab["wn_f"], np.int64)
    bad = lib().orc_dyson_g(nt, tab["nw_f"], _p(_c(tab["Twt_ff"])), _p(_c(tab["Ttw_ff"])), _p(_c(tab["T_beta"])),
                            _p(wn), ctypes.c_double(tab["beta"]), nk, no, _opt(H0), _p(F), _opt(S), _p(Sigma),
                            ctypes.c_double(mu), _p(G), _p(Dm))
    if bad:
        raise np.linalg.LinAlgError("oracle dyson_g: singular matrix at (n, k) index %d" % (bad - 1))
    return G, Dm


def nelec(tab, F, Sigma, mu, H0=None, S=None):
    """Electron number at mu (compute_Nelec, scf_common.cpp:10-42)."""
    F, Sigma = _c(F), _c(Sigma)
    nk, no = F.shape[-3], F.shape[-1]
    Sw = transform(tab["Twt_ff"], Sigma)
    wn = np.ascontiguousarray(tab["wn_f"], np.int64)
    return lib().orc_nelec(tab["nt"], tab["nw_f"], _p(_c(tab["Ttw_ff"])), _p(_c(tab["T_beta"])), _p(wn),
                           ctypes.c_double(tab["beta"]), nk, no, _opt(H0), _p(F), _opt(S), _p(_c(Sw)),
                           ctypes.c_double(mu))


def update_mu(tab, F, Sigma, nel_target, old_mu, mu_tol=1e-9, H0=None, S=None):
    """Chemical potential by bracketing and bisection (update_mu, scf_common.cpp:147-202)."""
    F, Sigma = _c(F), _c(Sigma)
    nk, no = F.shape[-3], F.shape[-1]
    wn = np.ascontiguousarray(tab["wn_f"], np.int64)
    nel = ctypes.c_double()
    mu = lib().orc_update_mu(tab["nt"], tab["nw_f"], _p(_c(tab["Twt_ff"])), _p(_c(tab["Ttw_ff"])),
                             _p(_c(tab["T_beta"])), _p(wn), ctypes.c_double(tab["beta"]), nk, no, _opt(H0), _p(F),
                             _opt(S), _p(Sigma), ctypes.c_double(nel_target), ctypes.c_double(old_mu),
                             ctypes.c_double(mu_tol), ctypes.byref(nel))
    return mu, nel.value


def hermitize(A):
    A = _c(A).copy()
    n = A.shape[-1]
    nb = int(np.prod(A.shape[:-2]))
    lib().orc_hermitize(ctypes.c_long(nb), n, _p(A))
    return A


def rpa_ec(tab, Piw, V, return_ew=False):
    Piw, V = _c(Piw), _c(V)
    nq, nmu = V.shape[0], V.shape[1]
    if Piw.shape != ((tab["nw_b"] + 1) // 2, nq, nmu, nmu) or tab["Ttw_bb"].shape != (tab["nt"], tab["nw_b"]):
        raise ValueError("rpa_ec: Pi_w %s does not match nw_half of the table and V %s" % (Piw.shape, V.shape))
    ew = np.zeros(tab["nw_b"], np.complex128)
    Ttw = _c(tab["Ttw_bb"])
    Tb = _c(tab["T_beta"])
    ec = lib().orc_rpa_ec(tab["nt"], tab["nw_b"], _p(Ttw), _p(Tb), nq, nmu, _p(Piw), _p(V), _p(ew))
    return (ec, ew) if return_ew else ec


def pade_eval(zf, Af, w):
    zf, Af, w = _c(zf), _c(Af), _c(w)
    out = np.zeros(w.shape, np.complex128)
    lib().orc_pade_eval(len(zf), _p(zf), _p(Af), len(w), _p(w), _p(out))
    return out


def solve_qp(tab, Sigma, C, Vhf, E0, mu=0.0, nfit=18, eta=1e-4, tol=1e-8):
    Sigma, C, Vhf = _c(Sigma), _c(C), _c(Vhf)
    E0 = np.ascontiguousarray(E0, np.float64)
    nk, no = E0.shape
    E = np.zeros((nk, no), np.float64)
    Twt = _c(tab["Twt_ff"])
    wn = np.ascontiguousarray(tab["wn_f"], np.int64)
    lib().orc_solve_qp(tab["nt"], tab["nw_f"], _p(Twt), _p(wn), ctypes.c_double(tab["beta"]), nk, no, _p(Sigma),
                       _p(C), _p(Vhf), _p(E0), ctypes.c_double(mu), int(nfit), ctypes.c_double(eta),
                       ctypes.c_double(tol), _p(E))
    return E


def gm_energy(tab, G, Sigma):
    G, Sigma = _c(G), _c(Sigma)
    nk, no = G.shape[2], G.shape[-1]
    return lib().orc_gm_energy(tab["nt"], tab["nw_f"], _p(_c(tab["Twt_ff"])), _p(_c(tab["Ttw_ff"])),
                               _p(_c(tab["T_beta"])), nk, no, _p(G), _p(Sigma))


def gw_iteration(tab, G, X, V, kgrid, with_ec=True):
    """One reference G0W0 iteration: update_w + gw_t::evaluate (scf_driver.cpp:119-122),
    plus the RPA energy (thc_rpa.icc) and hermitize (scf_driver.cpp:128)."""
    Pit = pi_rspace(G, X, kgrid)
    Piw = tau_to_w_phsym(tab, Pit)
    ec = rpa_ec(tab, Piw, V) if with_ec else None
    Ww = dyson(Piw, V)
    Wqt = w_to_qt(w_to_tau_phsym(tab, Ww))
    S = hermitize(sigma_rspace(G, X, Wqt, kgrid))
    return {"Pi_t": Pit, "Pi_w": Piw, "W_w": Ww, "W_qt": Wqt, "Sigma": S, "Ec": ec}


def eps_inv_head(tab, Wqt, chi_bar, qcart, volume, mode, q_max=0.8, fit_order=2):
    """(eps_tq (nt_half, nq), eps_t (nt_half)) of div_utils::eps_inv_head_t
    (g0_div_utils.hpp:55-111, eval_eps_inv_q :180-214, helpers :264-322).
    mode 0 ignore_g0 / 1 gygi: smallest nonzero |q|; 2 gygi_extrplt, 3 gygi_extrplt_2d."""
    nq = Wqt.shape[0]
    qa = np.sqrt(np.sum(np.asarray(qcart) ** 2, axis=1))
    # sum_P cbar_P sum_Q W_PQ conj(cbar_Q)  (nda::blas::dot does not conjugate)
    head = np.einsum("qp,qtpr,qr->tq", chi_bar, Wqt, chi_bar.conj())
    eps = head * (qa ** 2 / (4 * np.pi) * volume)[None, :]
    if nq == 1:
        mode = 0
    if mode in (2, 3):
        sel, seen = [], []
        for q in range(nq):
            r = qa[q]
            dup = any(abs(v - r) <= 1e-10 * max(v, r) for v in seen)  # equal lengths up to rounding
            if 0.0 < r <= q_max and (mode != 3 or qcart[q][2] == 0.0) and not dup:
                seen.append(r)
                sel.append(q)
        if len(sel) < 3:
            return eps, eps[:, sel[0]].copy()
        ew = tau_to_w_phsym(tab, np.ascontiguousarray(eps[:, sel]))
        A = np.array([[qa[q] ** j for j in range(fit_order + 1)] for q in sel])
        x = np.linalg.solve(A.T @ A, A.T @ ew.T)  # (fit_order+1, nw_half)
        e0 = w_to_tau_phsym(tab, np.ascontiguousarray(x[0][:, None]))[:, 0]
        return eps, e0
    if nq == 1:
        return eps, eps[:, 0].copy()
    cand = [q for q in range(nq) if qa[q] > 0.0]
    idx = min(cand, key=lambda q: (qa[q], q))
    return eps, eps[:, idx].copy()


def sigma_div_correction(Sigma, G, X, chi_head, madelung, eps_t, mode):
    """Sigma + the Gygi correction of gw_t::Sigma_div_correction (thc_gw.icc:446-526)."""
    nt, _, nk = Sigma.shape[:3]
    if mode == 0 or nk == 1:
        return Sigma.copy()
    nth = (nt + 1) // 2
    T = np.einsum("kpi,kpa,p->kia", X.conj(), X, chi_head.conj())
    out = Sigma.copy()
    for t in range(nt):
        tp = t if t < nth else nt - t - 1
        f = -madelung * eps_t[tp].real
        out[t, 0] += f * np.einsum("kia,kab,kjb->kij", T, G[t, 0], T.conj())
    return out
