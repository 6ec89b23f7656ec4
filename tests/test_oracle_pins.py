"""Pins of the FP64 CPU oracle (no GPU).

Each test checks an oracle function against data or mathematics that does not come from
the oracle itself: the reference's IAFT golden file, closed forms, the reference's second
algorithm for the same quantity, or numpy/LAPACK on small shapes.
"""
import os

import numpy as np
import pytest

from coqui_amd.system import make_system

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def small_system(kgrid=(2, 2, 1), no=6, nmu=10, sid=7):
    return make_system(kgrid, no, nmu, system_id=sid)


# ----------------------------------------------------------------- IR transforms
@pytest.mark.parametrize("prec,table,tol", [("high", "1e4.1e-15", 1e-11), ("medium", "1e4.1e-10", 1e-9)])
def test_iaft_golden_fixture(orc, prec, table, tol):
    """test_iaft.cpp:46-119: tau<->iw round trip and beta^- against the sparse_ir fixture."""
    z = np.load(os.path.join(GOLDEN, "iaft_si_kp222_%s.npz" % prec))
    tab = orc.load_ir(1000.0, None, table=table)
    Gt, Gw, Dm = z["Gt"], z["Gw"], z["Dm"]
    Gw_c = orc.transform(tab["Twt_ff"], Gt)
    Gt_c = orc.transform(tab["Ttw_ff"], Gw)
    # ARRAY_EQUAL(a, b, tol) of test_common.hpp: elementwise WithinRel(tol) || WithinAbs(tol)
    assert np.all((np.abs(Gw_c - Gw) <= tol) | (np.abs(Gw_c - Gw) <= tol * np.abs(Gw)))
    assert np.all((np.abs(Gt_c - Gt) <= tol) | (np.abs(Gt_c - Gt) <= tol * np.abs(Gt)))
    Dm_c = -orc.tau_to_beta(tab, Gt)
    assert np.all(np.abs(Dm_c - Dm) <= tol)


@pytest.mark.parametrize("table", ["1e2.1e-15", "1e3.1e-15", "1e4.1e-15", "1e5.1e-06"])
def test_iaft_identity(orc, table):
    """test_iaft.cpp:40-43: Ttw_ff Twt_ff = I (1e-10); tables rescaled as ir_driver.hpp:68-73."""
    tab = orc.load_ir(1000.0, None, table=table)
    eye = tab["Ttw_ff"] @ tab["Twt_ff"]
    assert np.max(np.abs(eye - np.eye(tab["nt"]))) < 1e-10
    eyeb = tab["Ttw_bb"] @ tab["Twt_bb"]
    assert np.max(np.abs(eyeb - np.eye(tab["nt"]))) < 1e-9


def single_pole_t(tab, e, beta):
    """g(tau) of qp_scf_common.cpp:86-93 (mu = 0) on the fermionic mesh, written directly."""
    t = tab["tau"]
    if e > 0:
        return -np.exp(-e * t) / (1 + np.exp(-e * beta))
    return -np.exp(e * (beta - t)) / (1 + np.exp(e * beta))


def test_iaft_single_pole(orc):
    """Closed form: a single pole g(tau) maps to 1/(iw_n - e) on the IR Matsubara mesh."""
    beta = 100.0
    tab = orc.load_ir(beta, None, table="1e3.1e-15")  # lambda = 1000 -> wmax = 10
    for e in (-0.7, -0.05, 0.3, 2.0):
        gt = single_pole_t(tab, e, beta)
        gw = orc.transform(tab["Twt_ff"], gt[:, None])[:, 0]
        iw = 1j * tab["wn_f"] * np.pi / beta
        assert np.max(np.abs(gw - 1.0 / (iw - e))) < 1e-10
        # tau -> beta^-: g(beta^-) = -f(e) (occupation), IAFT.icc:219-238
        gb = orc.tau_to_beta(tab, gt[:, None])[0]
        assert abs(gb + 1.0 / (1.0 + np.exp(beta * e))) < 1e-11


def test_ph_folding_matches_full_transform(orc):
    """IAFT.icc:49-73/104-128: for a bosonic function symmetric under tau -> beta - tau the
    folded half-mesh transforms equal the full-grid ones."""
    beta = 1000.0
    tab = orc.load_ir(beta, None, table="1e2.1e-15")
    nt, nth = tab["nt"], (tab["nt"] + 1) // 2
    t = tab["tau"]
    f = np.stack([np.exp(-a * t) + np.exp(-a * (beta - t)) for a in (0.01, 0.05, 0.2)], axis=1).astype(complex)
    f[:, 1] *= 1 + 0.5j
    full_w = orc.transform(tab["Twt_bb"], f)
    half_w = orc.tau_to_w_phsym(tab, f[:nth])
    nwb = tab["nw_b"]
    assert rel(half_w, full_w[nwb // 2:]) < 1e-12
    # and back: w_to_tau_PHsym of the half frequency data reproduces the tau data
    back = orc.w_to_tau_phsym(tab, full_w[nwb // 2:])
    assert rel(back, f[:nth]) < 1e-10


# ----------------------------------------------------------------- k <-> R
@pytest.mark.parametrize("kgrid", [(2, 2, 2), (3, 3, 3), (4, 4, 4), (2, 3, 1)])
def test_kR_coefficients(orc, kgrid):
    fRk, fkR = orc.kR_coefficients(kgrid)
    nk = int(np.prod(kgrid))
    assert np.max(np.abs(fkR @ fRk - np.eye(nk))) < 1e-13
    x = np.random.default_rng(1).standard_normal((nk, 3)) + 0j
    via_fft = np.fft.fftn(x.reshape(tuple(kgrid) + (3,)), axes=(0, 1, 2)).reshape(nk, 3) / nk
    assert np.max(np.abs(fRk @ x - via_fft)) < 1e-13


# ----------------------------------------------------------------- G and projections
def test_update_G_closed_form(orc):
    """qp_scf_common.cpp:82-118: G = C g C^H; G(0+) + G(beta-) -> -I, tau_to_beta(G) = -C f C^H."""
    s = small_system()
    beta = 1000.0
    tab = orc.load_ir(beta, None, table="1e3.1e-15")
    G = orc.update_G(tab, s.C, s.E)
    nk, no = s.E.shape
    for k in range(nk):
        for it in (0, 5, tab["nt"] - 1):
            g = np.array([single_pole_t({"tau": tab["tau"][it:it + 1]}, e, beta)[0] for e in s.E[k]])
            ref = (s.C[k] * g) @ s.C[k].conj().T
            assert np.max(np.abs(G[it, 0, k] - ref)) < 1e-13
    assert np.max(np.abs(G - np.conj(np.swapaxes(G, -1, -2)))) < 1e-14
    Gb = orc.tau_to_beta(tab, G)
    occ = 0.5 * (1.0 - np.tanh(0.5 * beta * s.E))
    Dm = np.einsum("kia,ka,kja->kij", s.C, occ, s.C.conj())
    assert np.max(np.abs(-Gb[0] - Dm)) < 1e-10


def test_projections_brute_force(orc):
    s = small_system()
    rng = np.random.default_rng(3)
    nk, no, nmu = s.nk, s.no, s.nmu
    O = rng.standard_normal((2, nk, no, no)) + 1j * rng.standard_normal((2, nk, no, no))
    P = orc.primary_to_aux(s.X, O)
    ref = np.einsum("kpa,nkab,kqb->nkpq", s.X, O, s.X.conj())
    assert rel(P, ref) < 1e-14
    Q = rng.standard_normal((2, nk, nmu, nmu)) + 1j * rng.standard_normal((2, nk, nmu, nmu))
    A = orc.aux_to_primary(s.X, Q)
    ref2 = np.einsum("kpa,nkpq,kqb->nkab", s.X.conj(), Q, s.X)
    assert rel(A, ref2) < 1e-14
    # X = [I; 0]: X O X^H is O padded with zeros
    Xi = np.zeros((nk, nmu, no), complex)
    Xi[:, :no, :no] = np.eye(no)
    Pi = orc.primary_to_aux(Xi, O)
    assert np.max(np.abs(Pi[:, :, :no, :no] - O)) == 0.0
    assert np.max(np.abs(Pi[:, :, no:, :])) == 0.0


# ----------------------------------------------------------------- chi0
def qminus(kgrid, q):
    n = np.array(kgrid)
    idx = np.array(np.unravel_index(q, kgrid))
    return int(np.ravel_multi_index(tuple((-idx) % n), kgrid))


@pytest.mark.parametrize("kgrid", [(2, 2, 2), (3, 1, 2)])
def test_pi_rspace_vs_kspace(orc, kgrid):
    """rpa_pi.icc: the R-space and k-space algorithms satisfy Pi^k_PQ(q) = Pi^R_QP(-q)."""
    s = make_system(kgrid, 5, 7, system_id=11)
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    G = orc.update_G(tab, s.C, s.E)
    PR = orc.pi_rspace(G, s.X, kgrid)
    PK = orc.pi_kspace(G, s.X, kgrid)
    nk = s.nk
    PRm = PR[:, [qminus(kgrid, q) for q in range(nk)]]
    assert rel(PK, np.swapaxes(PRm, -1, -2)) < 1e-12
    # Hermitian in PQ for Hermitian G
    assert rel(PR, np.conj(np.swapaxes(PR, -1, -2))) < 1e-12


def test_pi_brute_force_convolution(orc):
    """Pi(tau_i, q) = (sp/N_k) sum_k G_PQ(tau_i, k+q) conj(G_PQ(beta - tau_i, k)), sp = -2."""
    kgrid = (2, 2, 1)
    s = make_system(kgrid, 4, 5, system_id=12)
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    G = orc.update_G(tab, s.C, s.E)
    PR = orc.pi_rspace(G, s.X, kgrid)
    nt, nk = tab["nt"], s.nk
    GP = np.einsum("kpa,tkab,kqb->tkpq", s.X, G[:, 0], s.X.conj())
    ref = np.zeros_like(PR)
    for it in range((nt + 1) // 2):
        for q in range(nk):
            for k in range(nk):
                kq = np.ravel_multi_index(tuple((np.array(np.unravel_index(k, kgrid)) + np.array(np.unravel_index(q, kgrid))) % np.array(kgrid)), kgrid)
                ref[it, q] += -2.0 / nk * GP[it, kq] * np.conj(GP[nt - 1 - it, k])
    assert rel(PR, ref) < 1e-13


# ----------------------------------------------------------------- THC Hartree-Fock
def density_matrix(s, beta=1000.0, mu=0.0):
    f = 1.0 / (1.0 + np.exp(np.clip(beta * (s.E - mu), -700, 700)))
    return np.einsum("kia,ka,kja->kij", s.C, f, s.C.conj())


def test_thc_hf_gamma_closed_form(orc):
    """N_k = 1: F = X^H [diag(J) - D^P o V] X with D^P = X Dm X^H and J = 2 V diag(D^P)."""
    s = make_system((1, 1, 1), 6, 11, system_id=21)
    Dm = density_matrix(s)
    DP = np.einsum("pa,ab,qb->pq", s.X[0], Dm[0], s.X[0].conj())
    J = 2.0 * s.V[0] @ np.diag(DP)
    FP = np.diag(J) - DP * s.V[0]
    ref = np.einsum("pa,pq,qb->ab", s.X[0].conj(), FP, s.X[0])
    F = orc.thc_hf(Dm, s.X, s.V, (1, 1, 1))
    assert rel(F[0], ref) < 1e-13
    Fx = orc.thc_hf(Dm, s.X, s.V, (1, 1, 1), hartree=False)
    assert rel(Fx[0], np.einsum("pa,pq,qb->ab", s.X[0].conj(), -DP * s.V[0], s.X[0])) < 1e-13


@pytest.mark.parametrize("kgrid", [(2, 2, 1), (3, 1, 2)])
def test_thc_hf_brute_force_convolution(orc, kgrid):
    """Exchange F^P(k) = -(1/N_k) sum_q D^P(k - q) o V(q); Hartree J = 2 V(0) mean_k diag D^P(k)."""
    s = make_system(kgrid, 5, 7, system_id=22)
    Dm = density_matrix(s)
    nk = s.nk
    DP = np.einsum("kpa,kab,kqb->kpq", s.X, Dm, s.X.conj())
    n = np.array(kgrid)
    FP = np.zeros_like(DP)
    for k in range(nk):
        for q in range(nk):
            kmq = np.ravel_multi_index(tuple((np.array(np.unravel_index(k, kgrid)) - np.array(np.unravel_index(q, kgrid))) % n), kgrid)
            FP[k] += -DP[kmq] * s.V[q] / nk
    J = 2.0 * s.V[0] @ np.mean(np.diagonal(DP, axis1=1, axis2=2), axis=0)
    FP += np.diag(J)[None]
    ref = np.einsum("kpa,kpq,kqb->kab", s.X.conj(), FP, s.X)
    F = orc.thc_hf(Dm, s.X, s.V, kgrid)
    assert rel(F, ref) < 1e-12
    # Hermitian Dm and V_q give a Hermitian exchange potential (a sum of Hadamard products of
    # Hermitian matrices) with a negative exchange energy; the Hartree part is Hermitian only
    # for a real V(q = 0), which the synthetic V is not
    Fx = orc.thc_hf(Dm, s.X, s.V, kgrid, hartree=False)
    assert rel(Fx, np.conj(np.swapaxes(Fx, -1, -2))) < 1e-12
    assert np.einsum("kij,kji->", Dm, Fx).real < 0.0


# ----------------------------------------------------------------- Dyson for G, mu search
def fermi_sum(E, mu, beta=1000.0):
    """2 sum_k w_k sum_b f(e_b - mu), w_k = 1/N_k."""
    x = np.clip(beta * (E - mu), -700, 700)
    return 2.0 * np.sum(1.0 / (1.0 + np.exp(x))) / E.shape[0]


def test_dyson_g_free_particle(orc):
    """Sigma = 0, F = H(k): the Matsubara Dyson equation returns update_G's closed-form G(tau)
    and Dm = C f C^H (lambda = 1e4 keeps every |e| <= 1 Ha inside the IR window)."""
    s = make_system((2, 1, 1), 7, 3, system_id=31)
    tab = orc.load_ir(1000.0, None, table="1e4.1e-15")
    Sig = np.zeros((tab["nt"], 1, s.nk, s.no, s.no), complex)
    for mu in (0.0, 0.13):
        G, Dm = orc.dyson_g(tab, s.H, Sig, mu)
        G0 = orc.update_G(tab, s.C, s.E, mu)
        assert rel(G, G0) < 1e-9
        f = 1.0 / (1.0 + np.exp(np.clip(1000.0 * (s.E - mu), -700, 700)))
        assert rel(Dm, np.einsum("kia,ka,kja->kij", s.C, f, s.C.conj())) < 1e-9


def test_nelec_free_particle_and_mu(orc):
    """Sigma = 0: nelec(mu) = 2 sum_k w_k sum_b f(e_b - mu); update_mu finds the mu where the
    Fermi sum reaches the target."""
    s = make_system((2, 1, 1), 7, 3, system_id=32)
    tab = orc.load_ir(1000.0, None, table="1e4.1e-15")
    Sig = np.zeros((tab["nt"], 1, s.nk, s.no, s.no), complex)
    for mu in (-0.3, 0.0, 0.41):
        assert abs(orc.nelec(tab, s.H, Sig, mu) - fermi_sum(s.E, mu)) < 1e-8
    # a target between two levels: any mu in the gap gives it; check the found mu reproduces it
    target = fermi_sum(s.E, 0.5 * (np.sort(s.E.ravel())[6] + np.sort(s.E.ravel())[7]))
    mu, nel = orc.update_mu(tab, s.H, Sig, target, 0.0, mu_tol=1e-9)
    assert abs(nel - target) < 1e-9 and abs(fermi_sum(s.E, mu) - target) < 1e-7


def test_nelec_matches_eigen_spectra(orc):
    """compute_Nelec with the reference's eigenvalue spectra of H + Sigma(iw_n)
    (compute_eigenspectra, simple_dyson.cpp:130-172) equals the oracle's resolvent trace."""
    s = make_system((2, 1, 1), 6, 3, system_id=33)
    tab = orc.load_ir(1000.0, None, table="1e4.1e-15")
    nt, nw = tab["nt"], tab["nw_f"]
    # a causal self-energy: one pole per k at energy 0.3 with a Hermitian positive weight
    rng = np.random.default_rng(5)
    B = rng.standard_normal((s.nk, s.no, s.no)) + 1j * rng.standard_normal((s.nk, s.no, s.no))
    Wt = 0.02 * B @ np.conj(np.swapaxes(B, 1, 2)) / s.no
    tau, beta, ep = tab["tau"], 1000.0, 0.3
    g = -np.exp(-ep * tau) / (1.0 + np.exp(-ep * beta))
    Sig = (g[:, None, None, None] * Wt[None])[:, None]
    Sw = orc.transform(tab["Twt_ff"], Sig)[:, 0]
    for mu in (-0.1, 0.05):
        Xw = np.zeros(nw, complex)
        for n in range(nw):
            z = mu + 1j * tab["wn_f"][n] * np.pi / beta
            for k in range(s.nk):
                lam = np.linalg.eigvals(s.H[k] + Sw[n, k])
                Xw[n] += np.sum(1.0 / (z - lam)) / s.nk
        ref = -2.0 * (tab["T_beta"] @ (tab["Ttw_ff"] @ Xw)).real
        assert abs(orc.nelec(tab, s.H, Sig, mu) - ref) < 1e-9 * max(1.0, abs(ref))


# ----------------------------------------------------------------- Dyson
def test_dyson_closed_form_and_lapack(orc):
    rng = np.random.default_rng(5)
    # N_mu = 1: W = V/(1 - V Pi) - V
    Pi = rng.standard_normal((3, 2, 1, 1)) + 1j * rng.standard_normal((3, 2, 1, 1))
    V = np.abs(rng.standard_normal((2, 1, 1))) + 0j
    W = orc.dyson(Pi, V)
    ref = V[None] / (1 - V[None] * Pi) - V[None]
    assert np.max(np.abs(W - ref)) < 1e-14
    # small shapes vs numpy.linalg.inv (LAPACK getrf/getri), scr_coulomb_t.cpp:180-196
    s = make_system((2, 1, 1), 4, 12, system_id=13)
    Pi = -(rng.standard_normal((2, 2, 12, 12)) + 1j * rng.standard_normal((2, 2, 12, 12))) * 0.1
    W = orc.dyson(Pi, s.V)
    for n in range(2):
        for q in range(2):
            A = np.linalg.inv(np.eye(12) - s.V[q] @ Pi[n, q]) - np.eye(12)
            assert rel(W[n, q], A @ s.V[q]) < 1e-13


def test_dyson_resolvent_identity(orc):
    """For Hermitian positive definite V and Hermitian negative semidefinite Pi, the oracle's
    W (LU of I - V Pi) equals (V^-1 - Pi)^-1 - V, and its ln det term equals
    ln det V + ln det(V^-1 - Pi): the identities of the GPU resolvent form (flag 6 = 2).
    A transposed product (I - Pi V)^-1 V would differ at O(1)."""
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    s = make_system((2, 1, 1), 4, 10, system_id=16)
    rng = np.random.default_rng(9)
    nwh = (tab["nw_b"] + 1) // 2
    B = rng.standard_normal((nwh, 2, 10, 10)) + 1j * rng.standard_normal((nwh, 2, 10, 10))
    Pi = -(B @ np.conj(np.swapaxes(B, -1, -2))) * 0.05
    W = orc.dyson(Pi, s.V)
    _, ew = orc.rpa_ec(tab, Pi, s.V, return_ew=True)
    nwb = tab["nw_b"]
    ref = np.zeros(nwb, complex)
    for n in range(nwh):
        for q in range(2):
            Vi = np.linalg.inv(s.V[q])
            assert rel(W[n, q], np.linalg.inv(Vi - Pi[n, q]) - s.V[q]) < 1e-12
            assert rel(np.linalg.solve(np.eye(10) - Pi[n, q] @ s.V[q], s.V[q]) - s.V[q], W[n, q]) > 1e-3
            ld = np.linalg.slogdet(s.V[q])[1] + np.linalg.slogdet(Vi - Pi[n, q])[1]
            v = np.sum(np.conj(s.V[q]) * Pi[n, q]) + ld
            ref[nwb // 2 + n] += v
            if n:
                ref[nwb // 2 - n] += v
    assert np.max(np.abs(ew - ref)) < 1e-11


# ----------------------------------------------------------------- Sigma
@pytest.mark.parametrize("nt_table", ["1e2.1e-15", "1e5.1e-06"])
def test_sigma_rspace_vs_kspace(orc, nt_table):
    """thc_gw.icc: R-space (:70-213) and k-space (:342-443) self-energies agree; odd nt covered."""
    kgrid = (2, 2, 1)
    s = make_system(kgrid, 5, 6, system_id=14)
    tab = orc.load_ir(1000.0, None, table=nt_table)
    G = orc.update_G(tab, s.C, s.E)
    rng = np.random.default_rng(9)
    nth = (tab["nt"] + 1) // 2
    Wqt = rng.standard_normal((s.nk, nth, 6, 6)) + 1j * rng.standard_normal((s.nk, nth, 6, 6))
    SR = orc.sigma_rspace(G, s.X, Wqt, kgrid)
    SK = orc.sigma_kspace(G, s.X, Wqt, kgrid)
    assert rel(SR, SK) < 1e-12
    assert np.all(np.isfinite(SR)) and np.max(np.abs(SR)) > 0


# ----------------------------------------------------------------- Ec
def test_rpa_ec_closed_form_nmu1(orc):
    """N_mu = 1: Ec = T_beta . Ttw_bb e_w / (2 N_k), e_w(+-n) = sum_q [Pi V + ln(1 - Pi V)]."""
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    rng = np.random.default_rng(6)
    nq, nwh = 3, (tab["nw_b"] + 1) // 2
    Pi = -np.abs(rng.standard_normal((nwh, nq, 1, 1))) + 0j
    V = np.abs(rng.standard_normal((nq, 1, 1))) + 0j
    ec, ew = orc.rpa_ec(tab, Pi, V, return_ew=True)
    x = (Pi[:, :, 0, 0] * V[None, :, 0, 0]).real
    e_half = np.sum(x + np.log(1 - x), axis=1)
    nwb = tab["nw_b"]
    ew_ref = np.zeros(nwb, complex)
    for n in range(nwh):
        ew_ref[nwb // 2 + n] += e_half[n]
        if n:
            ew_ref[nwb // 2 - n] += e_half[n]
    assert np.max(np.abs(ew - ew_ref)) < 1e-13
    ec_ref = (tab["T_beta"] @ (tab["Ttw_bb"] @ ew_ref)).real / (2 * nq)
    assert abs(ec - ec_ref) < 1e-13 * max(1, abs(ec_ref))


def test_rpa_ec_logdet_vs_numpy(orc):
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    s = make_system((2, 1, 1), 4, 9, system_id=15)
    rng = np.random.default_rng(8)
    nwh = (tab["nw_b"] + 1) // 2
    B = rng.standard_normal((nwh, 2, 9, 9)) + 1j * rng.standard_normal((nwh, 2, 9, 9))
    Pi = -np.einsum("nqij,nqkj->nqik", B, B.conj()) * 0.05  # Hermitian negative semidefinite
    ec, ew = orc.rpa_ec(tab, Pi, s.V, return_ew=True)
    nwb = tab["nw_b"]
    ref = np.zeros(nwb, complex)
    for n in range(nwh):
        for q in range(2):
            A = Pi[n, q] @ s.V[q]
            sign, ld = np.linalg.slogdet(np.eye(9) - A)
            v = np.trace(A) + ld + np.log(abs(np.cos(np.angle(sign))))
            ref[nwb // 2 + n] += v
            if n:
                ref[nwb // 2 - n] += v
    assert np.max(np.abs(ew - ref)) < 1e-12


# ----------------------------------------------------------------- Pade / QP / GM
def test_pade_reproduces_rational(orc):
    """Thiele's continued fraction through N points reproduces a rational function of
    degree <= (N-1)/2 exactly (pade_t.hpp:176-192)."""
    beta = 1000.0
    z = 1j * (2 * np.arange(18) + 1) * np.pi / beta * 5
    poles, res = np.array([-0.3, 0.1, 0.45]), np.array([0.2, 0.5, 0.3])
    f = lambda w: np.sum(res[None, :] / (w[:, None] - poles[None, :]), axis=1)
    w = np.array([0.05 + 1e-4j, -0.2 + 1e-4j, 0.7 + 0.01j, 1j])
    out = orc.pade_eval(z, f(z), w)
    assert np.max(np.abs(out - f(w))) < 1e-9


def test_qp_single_pole_root(orc):
    """QP equation with a one-pole self-energy Sigma(w) = a/(w - p): the root of
    e - e0 - Re Sigma(e + i eta) = 0 is known in closed form (qp_scf_common.cpp:679-728)."""
    beta = 1000.0
    tab = orc.load_ir(beta, None, table="1e3.1e-15")
    a, p = 0.01, -0.8
    # Sigma(tau) of a single pole at p with weight a (fermionic): a * g_p(tau)
    sig_t = a * single_pole_t(tab, p, beta)
    nk, no = 1, 1
    Sigma = np.zeros((tab["nt"], 1, nk, no, no), complex)
    Sigma[:, 0, 0, 0, 0] = sig_t
    C = np.ones((1, 1, 1), complex)
    e0 = 0.2
    Vhf = np.full((1, 1, 1), e0, complex)
    E = orc.solve_qp(tab, Sigma, C, Vhf, np.array([[e0]]), mu=0.0, nfit=18, eta=1e-4, tol=1e-10)
    # e - e0 - a/(e - p) = 0  ->  e^2 - (e0 + p) e + e0 p - a = 0 ; root near e0
    disc = np.sqrt((e0 + p) ** 2 - 4 * (e0 * p - a))
    root = ((e0 + p) + disc) / 2
    assert abs(E[0, 0] - root) < 1e-6


def test_gm_energy_diagonal(orc):
    """eval_corr_energy with G = g(e) and Sigma = s * g(e') on one orbital:
    E = -sum_n Sigma(iw_n) G(iw_n)/beta through the IR tau->beta^- interpolation."""
    beta = 100.0
    tab = orc.load_ir(beta, None, table="1e3.1e-15")
    e1, e2, s = -0.3, 0.4, 0.05
    G = np.zeros((tab["nt"], 1, 1, 1, 1), complex)
    S = np.zeros_like(G)
    G[:, 0, 0, 0, 0] = single_pole_t(tab, e1, beta)
    S[:, 0, 0, 0, 0] = s * single_pole_t(tab, e2, beta)
    e = orc.gm_energy(tab, G, S)
    # (1/beta) sum_n s/((iw-e1)(iw-e2)) e^{iw 0+} = s (f(e1) - f(e2)) / (e1 - e2);
    # tau_to_beta(x) = -(1/beta) sum_n x(iw_n), and E = -0.5 * 2 * tau_to_beta(Sigma G)
    f = lambda x: 1.0 / (1.0 + np.exp(beta * x))
    ref = s * (f(e1) - f(e2)) / (e1 - e2)
    assert abs(e - ref) < 1e-9


# ---- the learned form's network (oracle/mlp_oracle.py) ----

def test_mlp_oracle_backward_finite_differences():
    """backward() against central differences of forward() in float64 (no BF16 rounding)."""
    from oracle import mlp_oracle as mo

    dims = (7, 11, 5)
    rng = np.random.default_rng(3)
    _, npar = mo.layout(*dims)
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((13, dims[0]))
    dY = rng.standard_normal((13, dims[2]))
    loss = lambda q: float(np.sum(mo.forward(q, dims, X, round_bf16=False)[0] * dY))
    _, cache = mo.forward(p, dims, X, round_bf16=False)
    g = mo.backward(cache, dims, dY, round_bf16=False)
    h = 1e-5
    for _ in range(6):  # random directions
        d = rng.standard_normal(npar)
        fd = (loss(p + h * d) - loss(p - h * d)) / (2 * h)
        assert abs(fd - g @ d) < 1e-7 * max(1.0, abs(fd))
    for idx in (0, npar // 3, npar // 2, npar - 1):  # single parameters in W0, W1, W2, b2
        e = np.zeros(npar)
        e[idx] = 1.0
        fd = (loss(p + h * e) - loss(p - h * e)) / (2 * h)
        assert abs(fd - g[idx]) < 1e-7 * max(1.0, abs(fd))


def test_mlp_oracle_bf16_rounding():
    """Round-to-nearest-even at the halfway points, and torch's bfloat16 conversion."""
    import torch

    from oracle import mlp_oracle as mo

    ulp = 2.0 ** -7
    cases = {1.0: 1.0, 1.0 + ulp / 2: 1.0, 1.0 + 1.5 * ulp: 1.0 + 2 * ulp, -(1.0 + ulp / 2): -1.0,
             3.0 + 2 * ulp * 0.49: 3.0}
    for x, want in cases.items():
        assert mo.bf16(np.array([x]))[0] == want
    x = np.random.default_rng(0).standard_normal(10000).astype(np.float32) * 100
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), ref)


def test_adam_oracle_closed_form():
    """First step moves every weight by lr g / (|g| + eps); a constant gradient keeps m_hat = g."""
    from oracle import mlp_oracle as mo

    g = np.array([0.3, -2.0, 1e-3])
    w, m, v = mo.adam(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, 0.01)
    assert np.allclose(w, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    for step in range(2, 6):
        w, m, v = mo.adam(w, m, v, g, step, 0.01)
    assert np.allclose(m / (1 - 0.9 ** 5), g, rtol=1e-12)
    assert np.allclose(np.sqrt(v / (1 - 0.999 ** 5)), np.abs(g), rtol=1e-12)


@pytest.mark.parametrize("kgrid", [(2, 2, 1), (3, 3, 1), (2, 3, 2)])
def test_time_reversal_relations(orc, kgrid):
    """A time-reversal symmetric system (H(-k) = conj H(k), X(-k) = conj X(k),
    V(-q) = conj V(q)) has Pi(-q) = conj Pi(q), W(-q) = conj W(q) and
    Sigma(-k) = conj Sigma(k): what lets the GPU path solve the irreducible q only."""
    from coqui_amd.system import minus_k

    s = make_system(kgrid, 6, 20, system_id=5, trev=True)
    mk = minus_k(kgrid)
    assert np.array_equal(s.V[mk], s.V.conj()) and np.array_equal(s.X[mk], s.X.conj())
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    G = orc.update_G(tab, s.C, s.E)
    Piw = orc.tau_to_w_phsym(tab, orc.pi_rspace(G, s.X, kgrid))
    W = orc.dyson(Piw, s.V)
    S = orc.gw_iteration(tab, G, s.X, s.V, kgrid, with_ec=False)["Sigma"]
    rel = lambda a, b: np.max(np.abs(a - b)) / np.max(np.abs(b))
    assert rel(Piw[:, mk], Piw.conj()) < 1e-12
    assert rel(W[:, mk], W.conj()) < 1e-12
    assert rel(S[:, :, mk], S.conj()) < 1e-12
    # a system drawn without the symmetry does not have it
    s2 = make_system(kgrid, 6, 20, system_id=5)
    assert rel(s2.V[mk], s2.V.conj()) > 0.1


# ---- long-wavelength (Gygi) correction ----

def _qcart(kgrid, L=10.0):
    """Cartesian q of a Gamma-centred grid, nearest images, cubic cell of side L (bohr)."""
    out = []
    for i in range(kgrid[0]):
        for j in range(kgrid[1]):
            for l in range(kgrid[2]):
                f = np.array([i / kgrid[0], j / kgrid[1], l / kgrid[2]])
                f = f - np.round(f)
                out.append(2 * np.pi * f / L)
    return np.array(out)


def test_eps_inv_head_hermitian_and_identity(orc):
    rng = np.random.default_rng(4)
    nq, nth, n = 4, 3, 7
    A = rng.standard_normal((nq, nth, n, n)) + 1j * rng.standard_normal((nq, nth, n, n))
    W = A + np.conj(np.swapaxes(A, 2, 3))
    cb = rng.standard_normal((nq, n)) + 1j * rng.standard_normal((nq, n))
    q = rng.uniform(0.1, 0.5, (nq, 3))
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    eps, _ = orc.eps_inv_head(tab, W, cb, q, 2.0, 1)
    assert np.max(np.abs(eps.imag)) < 1e-13 * np.max(np.abs(eps))
    # the conjugate on the other side would not be real: the pin discriminates
    wrong = np.einsum("qp,qtpr,qr->tq", cb.conj(), W, cb)
    assert np.max(np.abs(np.einsum("qp,qtpr,qr->tq", cb, W, cb).imag)) > 1e-3
    assert np.allclose(wrong.imag, 0, atol=1e-12 * np.max(np.abs(wrong)))  # also real, but differs:
    assert np.max(np.abs(wrong - eps / ((np.sum(q**2, 1) / (4 * np.pi) * 2.0)[None]))) > 1e-3
    # W = a I: eps = |q|^2/(4 pi) Vol a ||cbar||^2
    W2 = np.broadcast_to(3.0 * np.eye(n), (nq, nth, n, n)).astype(complex)
    eps2, _ = orc.eps_inv_head(tab, W2, cb, q, 2.0, 1)
    want = np.sum(q**2, 1) / (4 * np.pi) * 2.0 * 3.0 * np.sum(np.abs(cb) ** 2, 1)
    assert np.allclose(eps2, np.broadcast_to(want, (nth, nq)), rtol=1e-13)


def test_eps_inv_head_q0_rules(orc):
    """The fit reproduces a quadratic in |q| exactly (value at 0); the closest-q rule; the
    first-q rule below 3 points."""
    kgrid = (4, 4, 4)
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    nth = (tab["nt"] + 1) // 2
    q = _qcart(kgrid)
    qa = np.sqrt(np.sum(q**2, 1))
    nq, n = len(q), 3
    rng = np.random.default_rng(8)
    # a PH-symmetric tau profile: a single bosonic pole, sampled on the half grid
    x = tab["tau"][:nth] if "tau" in tab else np.linspace(0, 500, nth)
    a = np.cosh(0.002 * (x - 500.0))
    b, c = 0.3 * a, -0.7 * a
    target = a[:, None] + b[:, None] * qa[None] + c[:, None] * qa[None] ** 2
    fac = qa**2 / (4 * np.pi)
    cb = np.zeros((nq, n), complex)
    cb[:, 0] = 1.0
    W = np.zeros((nq, nth, n, n), complex)
    nz = qa > 0
    W[nz, :, 0, 0] = (target[:, nz] / fac[nz][None]).T
    eps, e0 = orc.eps_inv_head(tab, W, cb, q, 1.0, 2)
    assert np.allclose(eps[:, nz], target[:, nz], rtol=1e-12)
    assert np.max(np.abs(e0 - a)) < 1e-9 * np.max(np.abs(a))
    _, e1 = orc.eps_inv_head(tab, W, cb, q, 1.0, 1)
    assert np.allclose(e1, eps[:, int(np.argmin(np.where(nz, qa, np.inf)))])
    _, e2 = orc.eps_inv_head(tab, W, cb, q, 1.0, 2, q_max=0.2)  # one distinct |q| (0.157)
    first = [i for i in range(nq) if 0 < qa[i] <= 0.2][0]
    assert np.allclose(e2, eps[:, first])
    # 3x3x1: |q| of (1/3, 0) and (2/3 - 1, 0) differ in the last bit; they are one length,
    # so two distinct lengths remain and the first q is used, not a singular fit
    q3 = _qcart((3, 3, 1))
    W3 = np.zeros((9, nth, n, n), complex)
    W3[:, :, 0, 0] = 1.0
    eps3, e3 = orc.eps_inv_head(tab, W3, cb[:9], q3, 1.0, 2)
    assert np.allclose(e3, eps3[:, 1])


def test_sigma_div_correction_closed_form(orc):
    """X = [I; 0]: T = diag(conj h) and Delta_ij = f conj(h_i) G_ij h_j, f = -madelung Re eps(tau')."""
    rng = np.random.default_rng(6)
    nt, nk, no, nmu = 6, 3, 4, 7
    X = np.zeros((nk, nmu, no), complex)
    X[:, :no, :] = np.eye(no)
    h = rng.standard_normal(nmu) + 1j * rng.standard_normal(nmu)
    G = rng.standard_normal((nt, 1, nk, no, no)) + 1j * rng.standard_normal((nt, 1, nk, no, no))
    S = rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)
    eps_t = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    out = orc.sigma_div_correction(S, G, X, h, 1.7, eps_t, 1)
    for t in range(nt):
        tp = t if t < 3 else nt - t - 1
        f = -1.7 * eps_t[tp].real
        want = S[t, 0] + f * np.conj(h[:no])[None, :, None] * G[t, 0] * h[:no][None, None, :]
        assert np.allclose(out[t, 0], want, rtol=1e-13, atol=1e-14)
    assert np.array_equal(orc.sigma_div_correction(S, G, X, h, 1.7, eps_t, 0), S)


@pytest.mark.parametrize("kgrid", [(4, 4, 1), (3, 3, 1), (2, 2, 2)])
def test_c4_symmetry_relations(orc, kgrid):
    """A system symmetric under the fourfold rotation, acting on the THC points as the
    permutation pi_a, has Pi(S^a q) = Pi(q)[pi_a][:, pi_a] and the same for W: what lets the
    GPU path solve Dyson on the irreducible q only (thcgw_set_qsym)."""
    s = make_system(kgrid, 6, 20, system_id=7, c4=True)
    src, perm = s.qsym
    assert len(set(src.tolist())) < s.nk
    tab = orc.load_ir(1000.0, None, table="1e2.1e-15")
    G = orc.update_G(tab, s.C, s.E)
    Piw = orc.tau_to_w_phsym(tab, orc.pi_rspace(G, s.X, kgrid))
    W = orc.dyson(Piw, s.V)
    for A in (Piw, W):
        for q in range(s.nk):
            m = perm[q]
            B = A[:, src[q]][:, m][:, :, m]
            assert np.max(np.abs(A[:, q] - B)) < 1e-12 * np.max(np.abs(A))
    # a map that is off by one transposition breaks the relation
    q = int(np.nonzero(src != np.arange(s.nk))[0][0])
    m = perm[q].copy()
    m[[0, 1]] = m[[1, 0]]
    assert np.max(np.abs(W[:, q] - W[:, src[q]][:, m][:, :, m])) > 1e-3 * np.max(np.abs(W))
