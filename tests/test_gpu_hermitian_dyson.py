"""The Hermitian (two-Cholesky) form of the Dyson solve and RPA determinant (flag 6 = 1).

It computes the same W = (I - V Pi)^-1 V - V and ln det(I - V Pi) as the LU form of the
reference (scr_coulomb_t.cpp:180-196, thc_rpa.icc:141-163) when V is Hermitian positive
definite and Pi Hermitian negative semidefinite, and falls back to LU per chunk
otherwise.  Tolerance as for the plain FP64 path: 1e-10 relative to the oracle.
"""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, make_system

pytestmark = pytest.mark.gpu

TOL = 1e-10


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
    print("case: kgrid=%s no=%d nmu=%d table=%s sid=%d" % (kgrid, no, nmu, table, sid), flush=True)
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
    ctx.set_flag(6, 1)
    return dict(s=s, tab=tab, G=G, ref=ref, ctx=ctx)


def test_hermitian_dyson_nk8_nmu200(nk8_nmu200, torch):
    c = nk8_nmu200
    PiW = torch.from_numpy(c["ref"]["Pi_w"].copy()).cuda()
    c["ctx"].dyson_w(PiW, c["s"].V)
    assert relmax(to_np(PiW), c["ref"]["W_w"]) < TOL
    W = to_np(PiW)
    # W + V is mirrored from its lower triangle: W is Hermitian to the rounding of the input V
    assert relmax(W, np.conj(np.swapaxes(W, -1, -2))) < 1e-14


def test_hermitian_terms_and_rpa_ec(nk8_nmu200, torch):
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


def test_hermitian_gw_iteration_nk8_nmu200(nk8_nmu200, torch):
    c = nk8_nmu200
    S, W, ec = c["ctx"].gw_iteration(c["G"], c["s"].X, c["s"].V, with_ec=True, hermitize=True)
    assert relmax(to_np(W), c["ref"]["W_qt"]) < TOL
    assert relmax(to_np(S), c["ref"]["Sigma"]) < TOL
    assert abs(ec - c["ref"]["Ec"]) < TOL * abs(c["ref"]["Ec"])


@pytest.mark.parametrize("kgrid,table,no,nmu", [
    ((3, 3, 3), "1e2.1e-15", 9, 37),      # order not a multiple of the 32 leaf or 64 tile
    ((1, 1, 1), "1e2.1e-15", 5, 70),
    ((2, 2, 2), "1e2.1e-15", 12, 129),    # one past a tile boundary
])
def test_hermitian_iteration_shapes(orc, torch, kgrid, table, no, nmu):
    s, ia, tab, G = system(orc, kgrid, no, nmu, table, sid=5)
    ref = orc.gw_iteration(tab, G, s.X, s.V, s.kgrid, with_ec=True)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    ctx.set_flag(6, 1)
    S, W, ec = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
    assert relmax(to_np(W), ref["W_qt"]) < TOL
    assert relmax(to_np(S), ref["Sigma"]) < TOL
    assert abs(ec - ref["Ec"]) < TOL * abs(ref["Ec"])


def test_hermitian_large_order(torch):
    """N_mu = 1100: a deep recursion of the Cholesky and the triangular solve."""
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    nmu, nq = 1100, 1
    rng = np.random.default_rng(22)
    Z = (rng.standard_normal((nq, nmu, nmu)) + 1j * rng.standard_normal((nq, nmu, nmu))) / np.sqrt(2)
    V = Z @ np.conj(np.swapaxes(Z, 1, 2)) / nmu
    ctx = ThcGW((1, 1, 1), 4, nmu, ia)
    ctx.set_flag(6, 1)
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
        assert relmax(W[n, 0], ref) < 1e-10
        sign, lref = np.linalg.slogdet(A)
        assert abs(ld[n] - lref) < 1e-10 * abs(lref)
        assert abs(tr[n] - np.trace(V[0] @ Pi[n, 0])) < 1e-10 * abs(tr[n])
    # Pi = 0: M = I and W = L L^H - V, zero up to the rounding of the Cholesky factor of V
    assert np.max(np.abs(W[2:])) < 1e-13 * np.max(np.abs(V))


def test_hermitian_falls_back_to_lu(torch):
    """Pi = 2 I with V = I makes M = I - L^H Pi L = -I indefinite: those chunks run LU
    (W = (I - 2I)^-1 - I = -2I); the chunks with Pi = 0 stay Hermitian (W = 0)."""
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    nmu = 40
    ctx = ThcGW((1, 1, 1), 2, nmu, ia)
    ctx.set_flag(6, 1)
    ctx.set_flag(3, 1)
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
