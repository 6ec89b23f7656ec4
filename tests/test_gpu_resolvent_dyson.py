"""The resolvent form of the Dyson solve and RPA determinant (flag 6 = 2).

W + V = (I - V Pi)^-1 V = (V^-1 - Pi)^-1, the same W = (I - V Pi)^-1 V - V and
ln det(I - V Pi) = ln det V + ln det(V^-1 - Pi) as the LU form of the reference
(scr_coulomb_t.cpp:180-196, thc_rpa.icc:141-163) for Hermitian positive definite V and
Hermitian negative semidefinite Pi, with the per-chunk LU fallback of the Hermitian form.

Its rounding error grows with the condition number of V (V^-1 is formed explicitly): the
bound used here is 1e-10 + cond(V) eps relative to the largest element (observed about
0.01 cond(V) eps; 5e-10 at cond(V) = 2.5e8).  On the seeded systems below cond(V) stays
under 8e6 and an FP64 CPU restatement of the form is within 3.3e-11 of the oracle's W.
"""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, make_system

pytestmark = pytest.mark.gpu

TOL = 1e-10
EPS = np.finfo(float).eps


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def to_np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


def system(orc, kgrid, no, nmu, table, sid=0):
    print("case: resolvent kgrid=%s no=%d nmu=%d table=%s sid=%d" % (kgrid, no, nmu, table, sid), flush=True)
    s = make_system(kgrid, no, nmu, system_id=sid)
    ia = IAFT.from_table(1000.0, table)
    tab = orc.load_ir(1000.0, None, table=table)
    G = orc.update_G(tab, s.C, s.E)
    return s, ia, tab, G


@pytest.fixture(scope="module")
def nk8_nmu200(orc, torch):
    s, ia, tab, G = system(orc, (2, 2, 2), 26, 200, "1e2.1e-15")
    ref = orc.gw_iteration(tab, G, s.X, s.V, s.kgrid, with_ec=True)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    ctx.set_flag(6, 2)
    return dict(s=s, tab=tab, G=G, ref=ref, ctx=ctx)


def test_resolvent_dyson_nk8_nmu200(nk8_nmu200, torch):
    c = nk8_nmu200
    PiW = torch.from_numpy(c["ref"]["Pi_w"].copy()).cuda()
    c["ctx"].dyson_w(PiW, c["s"].V)
    assert relmax(to_np(PiW), c["ref"]["W_w"]) < TOL
    W = to_np(PiW)
    assert relmax(W, np.conj(np.swapaxes(W, -1, -2))) < 1e-14


def test_resolvent_chunking_bit_exact(nk8_nmu200, torch):
    """Each matrix's arithmetic does not depend on how the matrices are chunked: the
    automatic chunk (one chunk here), ragged chunks of 7 (fewer than the 8 q, so each V_q
    serves one matrix per chunk) and chunks of 20 (several matrices per V_q) give
    bit-identical W, Tr(V Pi) and ln det, also with the chunks pipelined over two streams
    (flag 11)."""
    c = nk8_nmu200
    outs = []
    for lanes, chunk in ((1, 0), (1, 7), (1, 20), (2, 0), (2, 7), (2, 20)):
        c["ctx"].set_flag(3, chunk)
        c["ctx"].set_flag(11, lanes)
        try:
            PiW = torch.from_numpy(c["ref"]["Pi_w"].copy()).cuda()
            tr, ld = c["ctx"].dyson_w(PiW, c["s"].V, with_terms=True)
        finally:
            c["ctx"].set_flag(3, 0)
            c["ctx"].set_flag(11, 1)
        outs.append((to_np(PiW), np.asarray(tr), np.asarray(ld)))
    for W, tr, ld in outs[1:]:
        assert np.array_equal(W, outs[0][0])
        assert np.array_equal(tr, outs[0][1]) and np.array_equal(ld, outs[0][2])


def test_resolvent_terms_and_rpa_ec(nk8_nmu200, torch):
    c = nk8_nmu200
    c["ctx"].set_flag(3, 7)  # ragged chunks
    try:
        PiW = torch.from_numpy(c["ref"]["Pi_w"].copy()).cuda()
        tr, ld = c["ctx"].dyson_w(PiW, c["s"].V, with_terms=True)
    finally:
        c["ctx"].set_flag(3, 0)
    assert relmax(to_np(PiW), c["ref"]["W_w"]) < TOL
    tab, nq = c["tab"], c["s"].nk
    nwb = tab["nw_b"]
    ew = np.zeros(nwb, complex)
    for g in range(len(tr)):
        n = g // nq
        ew[nwb // 2 + n] += tr[g] + ld[g]
        if n:
            ew[nwb // 2 - n] += tr[g] + ld[g]
    ec = (tab["T_beta"] @ (tab["Ttw_bb"] @ ew)).real / (2 * nq)
    assert abs(ec - c["ref"]["Ec"]) < TOL * abs(c["ref"]["Ec"])
    ec2 = c["ctx"].rpa_ec(c["ref"]["Pi_w"], c["s"].V)
    assert abs(ec2 - c["ref"]["Ec"]) < TOL * abs(c["ref"]["Ec"])


def test_resolvent_gw_iteration_nk8_nmu200(nk8_nmu200, torch):
    c = nk8_nmu200
    S, W, ec = c["ctx"].gw_iteration(c["G"], c["s"].X, c["s"].V, with_ec=True, hermitize=True)
    assert relmax(to_np(W), c["ref"]["W_qt"]) < TOL
    assert relmax(to_np(S), c["ref"]["Sigma"]) < TOL
    assert abs(ec - c["ref"]["Ec"]) < TOL * abs(c["ref"]["Ec"])


def test_resolvent_mixed_nk8_nmu200(nk8_nmu200, torch):
    """FP32 projections and k-grid transforms (flag 7 = 1) with the FP64 resolvent Dyson
    step: the mixed form's tolerances (W, Sigma 5e-5 relative; Ec 1e-6 Ha)."""
    c = nk8_nmu200
    c["ctx"].set_flag(7, 1)
    try:
        S, W, ec = c["ctx"].gw_iteration(c["G"], c["s"].X, c["s"].V, with_ec=True, hermitize=True)
    finally:
        c["ctx"].set_flag(7, 0)
    e_w, e_s, e_ec = relmax(to_np(W), c["ref"]["W_qt"]), relmax(to_np(S), c["ref"]["Sigma"]), abs(ec - c["ref"]["Ec"])
    print("mixed + resolvent: W %.2e Sigma %.2e Ec %.2e" % (e_w, e_s, e_ec))
    assert e_w < 5e-5 and e_s < 5e-5 and e_ec < 1e-6


@pytest.mark.parametrize("kgrid,table,no,nmu", [
    ((3, 3, 3), "1e2.1e-15", 9, 37),      # order not a multiple of the 32 leaf or 64 tile
    ((1, 1, 1), "1e2.1e-15", 5, 70),
    ((2, 2, 2), "1e2.1e-15", 12, 129),    # one past a tile boundary
])
@pytest.mark.parametrize("g3", [0, 1])  # flag 12: FP64 complex products by four or three real products
def test_resolvent_iteration_shapes(orc, torch, kgrid, table, no, nmu, g3):
    s, ia, tab, G = system(orc, kgrid, no, nmu, table, sid=5)
    ref = orc.gw_iteration(tab, G, s.X, s.V, s.kgrid, with_ec=True)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    ctx.set_flag(6, 2)
    try:
        ctx.set_flag(12, g3)
        S, W, ec = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
    finally:
        ctx.set_flag(12, 0)
    assert relmax(to_np(W), ref["W_qt"]) < TOL
    assert relmax(to_np(S), ref["Sigma"]) < TOL
    assert abs(ec - ref["Ec"]) < TOL * abs(ref["Ec"])


def test_resolvent_large_order(torch):
    """N_mu = 1100: deep recursions of the Cholesky factorisation and the triangular
    inverse.  Against numpy's LU solve per matrix, to 1e-10 + cond(V) eps relative (the
    explicit V^-1 and the final subtraction of V)."""
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    nmu, nq = 1100, 1
    print("case: resolvent large order nmu=%d" % nmu, flush=True)
    rng = np.random.default_rng(22)
    Z = (rng.standard_normal((nq, nmu, nmu)) + 1j * rng.standard_normal((nq, nmu, nmu))) / np.sqrt(2)
    V = Z @ np.conj(np.swapaxes(Z, 1, 2)) / nmu
    bound = 1e-10 + np.linalg.cond(V[0]) * EPS
    ctx = ThcGW((1, 1, 1), 4, nmu, ia)
    ctx.set_flag(6, 2)
    ctx.set_flag(3, 2)
    nwh = ia.nw_half
    B = rng.standard_normal((2, nq, nmu, nmu)) + 1j * rng.standard_normal((2, nq, nmu, nmu))
    Pi = np.zeros((nwh, nq, nmu, nmu), complex)
    Pi[:2] = -(B @ np.conj(np.swapaxes(B, -1, -2))) / nmu * 0.3
    PiW = torch.from_numpy(Pi).cuda()
    tr, ld = ctx.dyson_w(PiW, V, with_terms=True)
    W = to_np(PiW)
    for n in range(2):
        A = np.eye(nmu) - V[0] @ Pi[n, 0]
        ref = np.linalg.solve(A, V[0]) - V[0]
        e = relmax(W[n, 0], ref)
        print("n=%d: W %.2e (bound %.2e)" % (n, e, bound))
        assert e < bound
        sign, lref = np.linalg.slogdet(A)
        assert abs(ld[n] - lref) < 1e-10 * abs(lref)
        assert abs(tr[n] - np.trace(V[0] @ Pi[n, 0])) < 1e-10 * abs(tr[n])
    # Pi = 0: A = V^-1 and W = (V^-1)^-1 - V, zero up to the rounding of the two inversions
    assert np.max(np.abs(W[2:])) < bound * np.max(np.abs(V))


@pytest.mark.parametrize("lanes", [1, 2])
def test_resolvent_falls_back_to_lu(torch, lanes):
    """Pi = 2 I with V = I makes A = V^-1 - Pi = -I indefinite: those chunks run LU
    (W = (I - 2I)^-1 - I = -2I); the chunks with Pi = 0 stay in the resolvent form (W = 0).
    With two lanes (flag 11) the failing chunks fall on both streams."""
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    nmu = 40
    ctx = ThcGW((1, 1, 1), 2, nmu, ia)
    ctx.set_flag(6, 2)
    ctx.set_flag(3, 1)
    ctx.set_flag(11, lanes)
    V = np.eye(nmu, dtype=complex)[None]
    Pi = np.zeros((ia.nw_half, 1, nmu, nmu), complex)
    Pi[0] = 2.0 * np.eye(nmu)
    Pi[3] = 2.0 * np.eye(nmu)
    PiW = torch.from_numpy(Pi).cuda()
    tr, ld = ctx.dyson_w(PiW, V, with_terms=True)
    W = to_np(PiW)
    for n in range(ia.nw_half):
        expect = -2.0 * np.eye(nmu) if n in (0, 3) else np.zeros((nmu, nmu))
        assert np.max(np.abs(W[n, 0] - expect)) < 1e-14
    assert abs(tr[0] - 2.0 * nmu) < 1e-12 and abs(ld[0]) < 1e-12  # |det(-I)| = 1
    assert abs(tr[1]) == 0.0 and abs(ld[1]) == 0.0


def test_flag6_range():
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    ctx = ThcGW((1, 1, 1), 2, 8, ia)
    from coqui_amd import ThcgwError

    with pytest.raises(ThcgwError):
        ctx.set_flag(6, 3)


def test_hermitian_g_projections(orc, torch):
    """Flag 10: the projections of a Hermitian G form their lower tiles and mirror them, and
    Sigma's back-projection uses the lower half of Sigma_PQ.  The iteration equals the
    full-product one to rounding (1e-12 in FP64, 1e-5 with the FP32 projections) with every
    GEMM tile width, and G that is not Hermitian is refused."""
    from coqui_amd import ThcgwError

    s, ia, tab, G = system(orc, (2, 2, 2), 10, 200, "1e2.1e-15", sid=6)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    # every GEMM tile width: flag 5 (FP64) 0 = automatic, 3 = 128 x 128; flag 8 (FP32) 0 = 64, 2 = 128
    for prec, tile_flag, tile in ((0, 5, 0), (0, 5, 3), (1, 8, 0), (1, 8, 2)):
        ctx.set_flag(7, prec)
        ctx.set_flag(tile_flag, tile)
        ctx.set_flag(10, 0)
        S0, W0, e0 = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
        S0, W0 = to_np(S0), to_np(W0)
        ctx.set_flag(10, 1)
        S1, W1, e1 = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
        S1, W1 = to_np(S1), to_np(W1)
        ctx.set_flag(tile_flag, 0)
        print("prec %d tile %d: W %.1e Sigma %.1e Ec %.1e" % (prec, tile, relmax(W1, W0), relmax(S1, S0),
                                                              abs(e1 - e0)))
        tol = 1e-12 if prec == 0 else 1e-5
        assert relmax(W1, W0) < tol and relmax(S1, S0) < tol and abs(e1 - e0) < tol * abs(e0)
    ctx.set_flag(7, 0)
    Gbad = G.copy()
    Gbad[3, 0, 1, 0, 1] += 1e-3  # one element breaks G = G^H
    with pytest.raises(ThcgwError):
        ctx.gw_iteration(Gbad, s.X, s.V, with_ec=True)
    ctx.set_flag(10, 0)
    ctx.gw_iteration(Gbad, s.X, s.V, with_ec=True)  # accepted without the flag
