"""CPU port of the reference's FP64 algorithm on the host's BLAS/LAPACK (numpy/scipy, OpenBLAS),
used as the timed CPU baseline.

TEST INFRASTRUCTURE: only bench.py's cpu_baseline leg and tests/ use this module.

It restates the same steps as the oracle (oracle/thc_oracle.cpp) with the library calls the
reference makes through nda/SLATE, so that the CPU time reflects a vendor-BLAS build of the
reference rather than the oracle's own loops:

* projections O = X G X^H as two ZGEMMs per k (thc_solver_comm.hpp:363-429);
* k <-> R as dense ZGEMMs with the f_Rk / f_kR coefficients (rpa_pi.icc:115-148,
  thc_gw.icc:137-191; kpoint_utils.hpp:461-680);
* Hadamard products of chi0 (rpa_pi.icc:150-163, sp = -2) and Sigma (thc_gw.icc:161-166);
* Dyson: A = V Pi, A <- I - A, getrf + getri, A <- A - I, W = A V (scr_coulomb_t.cpp:171-199);
* RPA terms: Tr(Pi V) and ln |Re det(I - Pi V)| by getrf (thc_rpa.icc:141-163);
* tau <-> iW as ZGEMMs with the folded tables (IAFT.icc:49-73, 104-128).

tests/test_blas_port.py checks every function against the oracle.
"""
import numpy as np
import scipy.linalg as sla

from oracle import oracle


def _dft(kgrid):
    return oracle.kR_coefficients(kgrid)


def project(X, G):
    """X_k G_k X_k^H for G (nb, nk, no, no), X (nk, nmu, no) -> (nb, nk, nmu, nmu)."""
    A = np.matmul(X[None], G)
    return np.matmul(A, np.conj(np.swapaxes(X, -1, -2))[None])


def back_project(X, S):
    """X_k^H S_k X_k for S (nb, nk, nmu, nmu) -> (nb, nk, no, no)."""
    Xh = np.conj(np.swapaxes(X, -1, -2))
    return np.matmul(np.matmul(Xh[None], S), X[None])


def pi_pair(G_pair, X, kgrid):
    """Pi(tau_0, q) from G at tau_0 and beta - tau_0 (G_pair (2, 1, nk, no, no)) -> (nk, nmu, nmu)."""
    fRk, fkR = _dft(kgrid)
    nk, nmu = X.shape[0], X.shape[1]
    M = nmu * nmu
    P = project(X, G_pair[:, 0])  # (2, nk, nmu, nmu)
    Gp = (fRk @ P[0].reshape(nk, M))
    Gn = (fRk @ P[1].reshape(nk, M))
    H = -2.0 * Gp * np.conj(Gn)
    return (fkR @ H).reshape(nk, nmu, nmu)


def sigma_pair(G_pair, X, W_q, kgrid):
    """Sigma(tau_0) and Sigma(beta - tau_0) from one W(tau_0, q) slice W_q (nk, nmu, nmu) ->
    (2, nk, no, no)."""
    fRk, fkR = _dft(kgrid)
    nk, nmu = X.shape[0], X.shape[1]
    M = nmu * nmu
    P = project(X, G_pair[:, 0])
    WR = fRk @ W_q.reshape(nk, M)
    out = []
    for i in range(2):
        GR = fRk @ P[i].reshape(nk, M)
        S = fkR @ (-GR * WR)
        out.append(back_project(X, S.reshape(1, nk, nmu, nmu))[0])
    return np.stack(out)


def dyson(Piw, V):
    """W - V for Pi(iW) (nw, nq, nmu, nmu), V (nq, nmu, nmu), by getrf + getri."""
    nw, nq, n = Piw.shape[0], Piw.shape[1], Piw.shape[2]
    out = np.empty_like(Piw)
    eye = np.eye(n)
    for w in range(nw):
        for q in range(nq):
            A = eye - V[q] @ Piw[w, q]
            A = sla.inv(A, overwrite_a=True, check_finite=False) - eye
            out[w, q] = A @ V[q]
    return out


def rpa_terms(Piw, V):
    """(Tr(Pi V), ln |Re det(I - Pi V)|) per (w, q)."""
    nw, nq, n = Piw.shape[0], Piw.shape[1], Piw.shape[2]
    tr = np.empty((nw, nq), np.complex128)
    ld = np.empty((nw, nq))
    eye = np.eye(n)
    for w in range(nw):
        for q in range(nq):
            A = Piw[w, q] @ V[q]
            tr[w, q] = np.trace(A)
            lu, piv = sla.lu_factor(eye - A, overwrite_a=True, check_finite=False)
            # ln|Re det| = sum ln|u_ii| + ln|Re(phase)|, phase = sign(perm) prod u_ii/|u_ii|
            # (the product of the diagonal itself overflows at N_mu in the thousands)
            d = np.diag(lu)
            sign = -1.0 if np.count_nonzero(piv != np.arange(n)) % 2 else 1.0
            ld[w, q] = np.sum(np.log(np.abs(d))) + np.log(abs((sign * np.prod(d / np.abs(d))).real))
    return tr, ld


def fold_tables(tab):
    """Folded T+ (nw_half, nt_half) and T- (nt_half, nw_half) of IAFT.icc:62-70 and :117-125."""
    T, Ti = tab["Twt_bb"], tab["Ttw_bb"]
    nwb, nt = T.shape
    nth, nwh = (nt + 1) // 2, (nwb + 1) // 2
    Tp = np.zeros((nwh, nth), complex)
    Tm = np.zeros((nth, nwh), complex)
    for n in range(nwh):
        for i in range(nth):
            j = nt - 1 - i
            Tp[n, i] = T[nwb // 2 + n, i] + (T[nwb // 2 + n, j] if j != i else 0)
    for i in range(nth):
        for n in range(nwh):
            Tm[i, n] = Ti[i, nwb // 2 + n] + (Ti[i, nwb // 2 - n] if n != 0 else 0)
    return Tp, Tm


def tau_to_w(Tp, Xt):
    return (Tp @ Xt.reshape(Xt.shape[0], -1)).reshape((Tp.shape[0],) + Xt.shape[1:])


def w_to_tau(Tm, Xw):
    return (Tm @ Xw.reshape(Xw.shape[0], -1)).reshape((Tm.shape[0],) + Xw.shape[1:])
