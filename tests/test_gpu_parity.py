"""Parity of the HIP path (through the C ABI) with the FP64 CPU oracle.

Tolerance: the plain FP64 GPU path must match the oracle to 1e-10 relative on chi0, W,
Sigma and the RPA energy (relative to the largest element of the oracle result).
"""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, ThcgwError, make_system

pytestmark = pytest.mark.gpu

TOL = 1e-10


def relmax(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def to_np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


def setup(orc, kgrid, no, nmu, table, sid=0, beta=1000.0):
    print("case: kgrid=%s no=%d nmu=%d table=%s sid=%d" % (kgrid, no, nmu, table, sid), flush=True)
    s = make_system(kgrid, no, nmu, system_id=sid)
    ia = IAFT.from_table(beta, table)
    tab = orc.load_ir(beta, None, table=table)
    G = orc.update_G(tab, s.C, s.E)
    return s, ia, tab, G


@pytest.fixture(scope="module")
def nk8_nmu200(orc, torch):
    """The N_k = 8 (2x2x2), N_orb = 26, N_mu = 200, N_tau = 40, beta = 1000 system."""
    s, ia, tab, G = setup(orc, (2, 2, 2), 26, 200, "1e2.1e-15")
    ref = orc.gw_iteration(tab, G, s.X, s.V, s.kgrid, with_ec=True)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    return dict(s=s, ia=ia, tab=tab, G=G, ref=ref, ctx=ctx)


def test_update_G(nk8_nmu200, torch):
    c = nk8_nmu200
    G = c["ctx"].update_G(c["s"].C, c["s"].E, 0.0)
    torch.cuda.synchronize()
    assert relmax(to_np(G), c["G"]) < 1e-13


def test_projections(nk8_nmu200, orc, torch):
    c, s = nk8_nmu200, nk8_nmu200["s"]
    rng = np.random.default_rng(1)
    O = rng.standard_normal((3, s.nk, s.no, s.no)) + 1j * rng.standard_normal((3, s.nk, s.no, s.no))
    P = c["ctx"].primary_to_aux(s.X, O)
    assert relmax(to_np(P), orc.primary_to_aux(s.X, O)) < 1e-13
    Q = rng.standard_normal((2, s.nk, s.nmu, s.nmu)) + 1j * rng.standard_normal((2, s.nk, s.nmu, s.nmu))
    A = c["ctx"].aux_to_primary(s.X, Q)
    assert relmax(to_np(A), orc.aux_to_primary(s.X, Q)) < 1e-13


def test_pi_nk8_nmu200(nk8_nmu200, torch):
    c = nk8_nmu200
    P = c["ctx"].pi(c["G"], c["s"].X)
    assert relmax(to_np(P), c["ref"]["Pi_t"]) < TOL


def test_tau_to_w_and_back(nk8_nmu200, orc, torch):
    c = nk8_nmu200
    Pw = c["ctx"].tau_to_w(c["ref"]["Pi_t"])
    assert relmax(to_np(Pw), c["ref"]["Pi_w"]) < TOL
    Wqt = c["ctx"].w_to_tau(c["ref"]["W_w"])
    assert relmax(to_np(Wqt), c["ref"]["W_qt"]) < TOL


def test_dyson_nk8_nmu200(nk8_nmu200, torch):
    c = nk8_nmu200
    PiW = torch.from_numpy(c["ref"]["Pi_w"].copy()).cuda()
    c["ctx"].dyson_w(PiW, c["s"].V)
    assert relmax(to_np(PiW), c["ref"]["W_w"]) < TOL


def test_dyson_chunked_and_terms(nk8_nmu200, orc, torch):
    """Small Dyson batches (chunk of 5 matrices) give the same W; Tr and ln det terms
    reproduce the oracle's RPA energy."""
    c = nk8_nmu200
    c["ctx"].set_flag(3, 5)
    try:
        PiW = torch.from_numpy(c["ref"]["Pi_w"].copy()).cuda()
        tr, ld = c["ctx"].dyson_w(PiW, c["s"].V, with_terms=True)
    finally:
        c["ctx"].set_flag(3, 0)
    assert relmax(to_np(PiW), c["ref"]["W_w"]) < TOL
    tab = c["tab"]
    nwb, nq = tab["nw_b"], c["s"].nk
    ew = np.zeros(nwb, complex)
    for g in range(len(tr)):
        n = g // nq
        ew[nwb // 2 + n] += tr[g] + ld[g]
        if n:
            ew[nwb // 2 - n] += tr[g] + ld[g]
    ec = (tab["T_beta"] @ (tab["Ttw_bb"] @ ew)).real / (2 * nq)
    assert abs(ec - c["ref"]["Ec"]) < TOL * abs(c["ref"]["Ec"])


def test_sigma_nk8_nmu200(nk8_nmu200, orc, torch):
    c = nk8_nmu200
    S = c["ctx"].sigma(c["G"], c["ref"]["W_qt"], c["s"].X, hermitize=True)
    assert relmax(to_np(S), c["ref"]["Sigma"]) < TOL
    S0 = c["ctx"].sigma(c["G"], c["ref"]["W_qt"], c["s"].X, hermitize=False)
    ref0 = orc.sigma_rspace(c["G"], c["s"].X, c["ref"]["W_qt"], c["s"].kgrid)
    assert relmax(to_np(S0), ref0) < TOL


def test_rpa_ec_nk8_nmu200(nk8_nmu200, torch):
    c = nk8_nmu200
    ec = c["ctx"].rpa_ec(c["ref"]["Pi_w"], c["s"].V)
    assert abs(ec - c["ref"]["Ec"]) < TOL * abs(c["ref"]["Ec"])


def test_gw_iteration_nk8_nmu200(nk8_nmu200, torch):
    c = nk8_nmu200
    S, W, ec = c["ctx"].gw_iteration(c["G"], c["s"].X, c["s"].V, with_ec=True, hermitize=True)
    assert relmax(to_np(S), c["ref"]["Sigma"]) < TOL
    assert relmax(to_np(W), c["ref"]["W_qt"]) < TOL
    assert abs(ec - c["ref"]["Ec"]) < TOL * abs(c["ref"]["Ec"])
    print("Ec gpu %.15e oracle %.15e" % (ec, c["ref"]["Ec"]))


def test_qp_and_gm_nk8_nmu200(nk8_nmu200, orc, torch):
    c, s = nk8_nmu200, nk8_nmu200["s"]
    H = s.H
    Sig = c["ref"]["Sigma"]
    E_gpu = c["ctx"].solve_qp(Sig, s.C, H, s.E, mu=0.0, nfit=18, eta=1e-4, tol=1e-8)
    E_orc = orc.solve_qp(c["tab"], Sig, s.C, H, s.E, mu=0.0, nfit=18, eta=1e-4, tol=1e-8)
    # the bisection stops at |residual| <= 1e-8 Ha; both sides run the same algorithm
    assert np.max(np.abs(E_gpu - E_orc)) < 1e-6
    e_gpu = c["ctx"].gm_energy(c["G"], Sig)
    e_orc = orc.gm_energy(c["tab"], c["G"], Sig)
    assert abs(e_gpu - e_orc) < TOL * max(1.0, abs(e_orc))


@pytest.mark.parametrize("kgrid,table,no,nmu", [
    ((3, 3, 3), "1e2.1e-15", 9, 37),      # N_k = 27 (radix-3), nmu not a multiple of 64
    ((4, 4, 4), "1e5.1e-06", 7, 24),      # N_k = 64 (lane-split DFT), odd nt = 61
    ((1, 1, 1), "1e2.1e-15", 5, 70),      # Gamma only
    ((2, 3, 1), "1e2.1e-15", 6, 20),      # non-cubic grid: generic transform kernel
    ((2, 2, 2), "1e3.1e-15", 6, 24),      # 72 tau points: tau <-> iW MFMA kernel with 3 row blocks
    ((2, 2, 2), "1e3.5e-18", 6, 24),      # 80 tau points: iW -> tau with 41 frequencies (11 K blocks)
    ((1, 1, 1), "1e4.1e-15", 5, 20),      # 104 tau points: tau <-> iW by the GEMM (T has 53 rows)
])
@pytest.mark.parametrize("g3", [0, 1])  # flag 12: FP64 complex products by four or three real products
def test_iteration_shapes(orc, torch, kgrid, table, no, nmu, g3):
    s, ia, tab, G = setup(orc, kgrid, no, nmu, table, sid=3)
    ref = orc.gw_iteration(tab, G, s.X, s.V, s.kgrid, with_ec=True)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    try:
        ctx.set_flag(12, g3)
        S, W, ec = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
    finally:
        ctx.set_flag(12, 0)
    assert relmax(to_np(W), ref["W_qt"]) < TOL
    assert relmax(to_np(S), ref["Sigma"]) < TOL
    assert abs(ec - ref["Ec"]) < TOL * abs(ref["Ec"])


def test_generic_kfft_matches_specialised(orc, torch):
    s, ia, tab, G = setup(orc, (2, 2, 2), 6, 30, "1e2.1e-15", sid=4)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    P1 = to_np(ctx.pi(G, s.X))
    ctx.set_flag(1, 1)
    P2 = to_np(ctx.pi(G, s.X))
    ctx.set_flag(1, 0)
    assert relmax(P2, P1) < 1e-13
    assert relmax(P1, orc.pi_rspace(G, s.X, s.kgrid)) < TOL


def test_dyson_large_order(orc, torch):
    """N_mu = 1100 (> 1024) takes the two-columns-per-thread LU panel path."""
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    nmu, nq = 1100, 1
    rng = np.random.default_rng(21)
    Z = (rng.standard_normal((nq, nmu, nmu)) + 1j * rng.standard_normal((nq, nmu, nmu))) / np.sqrt(2)
    V = Z @ np.conj(np.swapaxes(Z, 1, 2)) / nmu
    ctx = ThcGW((1, 1, 1), 4, nmu, ia)
    nwh = ia.nw_half
    B = rng.standard_normal((2, nq, nmu, nmu)) + 1j * rng.standard_normal((2, nq, nmu, nmu))
    Pi = np.zeros((nwh, nq, nmu, nmu), complex)
    Pi[:2] = -(B @ np.conj(np.swapaxes(B, -1, -2))) / nmu * 0.3
    ctx.set_flag(3, 2)
    PiW = torch.from_numpy(Pi).cuda()
    ctx.dyson_w(PiW, V)
    W = to_np(PiW)
    for n in range(2):
        ref = np.linalg.solve(np.eye(nmu) - V[0] @ Pi[n, 0], V[0]) - V[0]
        assert relmax(W[n, 0], ref) < 1e-10
    # zero Pi gives W - V = 0
    assert np.max(np.abs(W[2:])) == 0.0


def test_dyson_singular_reports_error(torch):
    """I - V Pi = 0 exactly: the LU reports a zero pivot and the call returns SINGULAR."""
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    nmu = 8
    ctx = ThcGW((1, 1, 1), 2, nmu, ia)
    V = np.eye(nmu, dtype=complex)[None]
    Pi = np.zeros((ia.nw_half, 1, nmu, nmu), complex)
    Pi[:] = np.eye(nmu)
    with pytest.raises(ThcgwError) as e:
        ctx.dyson_w(torch.from_numpy(Pi).cuda(), V)
    assert e.value.code == 3


def test_create_rejects_non_ph_symmetric_mesh(torch):
    """scr_coulomb_t.cpp:26-34: a tau mesh that is not symmetric about beta/2 is refused."""
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    ia.tau_mesh = ia.tau_mesh.copy()
    ia.tau_mesh[-1] = 0.5  # |x_0| - |x_{nt-1}| > 1e-6
    with pytest.raises(ThcgwError) as e:
        ThcGW((1, 1, 1), 2, 4, ia)
    assert "particle-hole" in str(e.value)


@pytest.mark.parametrize("kgrid,no,nmu", [((2, 2, 2), 26, 200), ((4, 4, 4), 7, 24), ((3, 1, 2), 6, 20)])
def test_kspace_variants(orc, torch, kgrid, no, nmu):
    """k-space chi0 and Sigma (rpa_pi.icc:182-336, thc_gw.icc:342-443) against the oracle's
    k-space restatement."""
    s, ia, tab, G = setup(orc, kgrid, no, nmu, "1e2.1e-15", sid=9)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    PK = to_np(ctx.pi_kspace(G, s.X))
    assert relmax(PK, orc.pi_kspace(G, s.X, s.kgrid)) < TOL
    Wqt = orc.gw_iteration(tab, G, s.X, s.V, s.kgrid, with_ec=False)["W_qt"]
    SK = to_np(ctx.sigma_kspace(G, Wqt, s.X, hermitize=False))
    assert relmax(SK, orc.sigma_kspace(G, s.X, Wqt, s.kgrid)) < TOL


@pytest.mark.parametrize("tile", [1, 2, 3])
def test_fp64_tile_shapes(orc, tile):
    """Every complex-FP64 GEMM tile shape (flag 5) gives the oracle's iteration (1e-10; the
    resolvent form 1e-10 + cond(V) eps)."""
    import torch

    kgrid, no, nmu, table = (2, 2, 2), 10, 150, "1e2.1e-15"
    print("case: tile shape %d" % tile, flush=True)
    s = make_system(kgrid, no, nmu, system_id=14)
    tab = orc.load_ir(1000.0, None, table=table)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, table))
    G = ctx.update_G(s.C, s.E, 0.0)
    ref = orc.gw_iteration(tab, orc.update_G(tab, s.C, s.E), s.X, s.V, kgrid, with_ec=True)
    try:
        ctx.set_flag(5, tile)
        for f6 in (0, 1, 2):
            ctx.set_flag(6, f6)
            S, W, ec = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
            torch.cuda.synchronize()
            # the resolvent form (flag 6 = 2) rounds with cond(V): 1e-10 + cond(V) eps
            to = 1e-10 if f6 < 2 else 1e-10 + max(np.linalg.cond(v) for v in s.V) * np.finfo(float).eps
            assert np.max(np.abs(S.cpu().numpy() - ref["Sigma"])) < to * np.max(np.abs(ref["Sigma"]))
            assert np.max(np.abs(W.cpu().numpy() - ref["W_qt"])) < to * np.max(np.abs(ref["W_qt"]))
            assert abs(ec - ref["Ec"]) < 1e-10 * abs(ref["Ec"])
    finally:
        ctx.set_flag(5, 0)


def test_fp64_three_products(orc):
    """Complex GEMMs by three real MFMA products (flag 12): in FP64 (bit 0) they give the
    oracle's iteration in every Dyson form (1e-10; the resolvent form 1e-10 + cond(V) eps).
    In the mixed forms the result stays within the mixed tolerances (W, Sigma 5e-5
    relative, Ec 1e-6 Ha) also with the FP32 products on three (bits 1, 2, 3), and with the
    FP64 ones only it differs from the four-product result by less than 10% of the mixed
    form's own error."""
    import torch

    kgrid, no, nmu, table = (2, 2, 2), 10, 150, "1e2.1e-15"
    print("case: three-product FP64 GEMMs", flush=True)
    s = make_system(kgrid, no, nmu, system_id=15)
    tab = orc.load_ir(1000.0, None, table=table)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, table))
    G = ctx.update_G(s.C, s.E, 0.0)
    ref = orc.gw_iteration(tab, orc.update_G(tab, s.C, s.E), s.X, s.V, kgrid, with_ec=True)
    try:
        ctx.set_flag(12, 1)
        for f6 in (0, 1, 2):
            ctx.set_flag(6, f6)
            S, W, ec = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
            torch.cuda.synchronize()
            to = 1e-10 if f6 < 2 else 1e-10 + max(np.linalg.cond(v) for v in s.V) * np.finfo(float).eps
            print("FP64 Dyson form %d: W %.2e Sigma %.2e" % (f6, relmax(W.cpu().numpy(), ref["W_qt"]),
                                                              relmax(S.cpu().numpy(), ref["Sigma"])))
            assert np.max(np.abs(S.cpu().numpy() - ref["Sigma"])) < to * np.max(np.abs(ref["Sigma"]))
            assert np.max(np.abs(W.cpu().numpy() - ref["W_qt"])) < to * np.max(np.abs(ref["W_qt"]))
            assert abs(ec - ref["Ec"]) < 1e-10 * abs(ref["Ec"])
        ctx.set_flag(7, 1)
        for f6 in (1, 2):
            ctx.set_flag(6, f6)
            ctx.set_flag(10, int(f6 == 2))
            out = {}
            for g3 in (0, 1, 5, 7, 13):
                ctx.set_flag(12, g3)
                S, W, ec = ctx.gw_iteration(G, s.X, s.V, with_ec=True, hermitize=True)
                out[g3] = (S.cpu().numpy(), W.cpu().numpy(), ec)
                e_w, e_s = relmax(out[g3][1], ref["W_qt"]), relmax(out[g3][0], ref["Sigma"])
                print("mixed, Dyson form %d, flag 12 = %d: W %.2e Sigma %.2e Ec %.2e"
                      % (f6, g3, e_w, e_s, abs(ec - ref["Ec"])))
                assert e_w < 5e-5 and e_s < 5e-5 and abs(ec - ref["Ec"]) < 1e-6
            d_w, d_s = relmax(out[1][1], out[0][1]), relmax(out[1][0], out[0][0])
            print("three vs four FP64 products: W %.2e Sigma %.2e" % (d_w, d_s))
            # Sigma takes W(tau) through FP32 k-grid transforms: where the two W differ by a
            # few FP64 ulps, an FP32 rounding of W can flip by one FP32 ulp
            assert d_w < 0.1 * relmax(out[0][1], ref["W_qt"]) and d_s < 0.1 * relmax(out[0][0], ref["Sigma"])
    finally:
        ctx.set_flag(12, 0)
        ctx.close()
