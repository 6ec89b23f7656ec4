"""Mixed-precision forms (flag 7 = 1: FP32 MFMA products; flag 7 = 2: split-BF16 MFMA
products) against the FP64 oracle.

Reduced-precision products (projections, k-grid transforms with the Hadamard products,
the triangular products of the Hermitian Dyson form) with FP64 tau <-> iW transforms,
Cholesky factorisations and ln det.  Tolerances (stated per quantity, relative to the largest
element of the oracle result unless noted):

  chi0(tau)  5e-6      W(tau)  5e-5      Sigma  5e-5
  RPA Ec     1e-6 Ha absolute
  QP energies 1 meV absolute (3.6749e-5 Ha)

On the lambda = 1e5 IR table the iW -> tau transform amplifies the FP32 rounding of
W(iW) about 30 times more than on the lambda = 1e2 table; W(tau) is held to 2e-4 there.
"""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, make_system

pytestmark = pytest.mark.gpu

HA_PER_MEV = 1.0 / 27211.386245988
TOL_PI, TOL_W, TOL_S, TOL_EC, TOL_QP = 5e-6, 5e-5, 5e-5, 1e-6, 1.0 * HA_PER_MEV


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
    return dict(s=s, tab=tab, G=G, ref=ref, ctx=ctx)


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_stages_nk8_nmu200(nk8_nmu200, torch, prec):
    c, s = nk8_nmu200, nk8_nmu200["s"]
    ctx, ref = c["ctx"], c["ref"]
    ctx.set_flag(7, prec)
    e_pi = relmax(to_np(ctx.pi(c["G"], s.X)), ref["Pi_t"])
    PiW = torch.from_numpy(ref["Pi_w"].copy()).cuda()
    ctx.dyson_w(PiW, s.V)
    e_w = relmax(to_np(PiW), ref["W_w"])
    e_s = relmax(to_np(ctx.sigma(c["G"], ref["W_qt"], s.X, hermitize=True)), ref["Sigma"])
    print("mixed(%d) stage errors: chi0 %.2e  W(iW) %.2e  Sigma %.2e" % (prec, e_pi, e_w, e_s))
    assert e_pi < TOL_PI and e_w < TOL_W and e_s < TOL_S


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_iteration_nk8_nmu200(nk8_nmu200, orc, torch, prec):
    c, s = nk8_nmu200, nk8_nmu200["s"]
    c["ctx"].set_flag(7, prec)
    S, W, ec = c["ctx"].gw_iteration(c["G"], s.X, s.V, with_ec=True, hermitize=True)
    S, W = to_np(S), to_np(W)
    e_w, e_s, e_ec = relmax(W, c["ref"]["W_qt"]), relmax(S, c["ref"]["Sigma"]), abs(ec - c["ref"]["Ec"])
    E_mix = orc.solve_qp(c["tab"], S, s.C, s.H, s.E)
    E_ref = orc.solve_qp(c["tab"], c["ref"]["Sigma"], s.C, s.H, s.E)
    e_qp = float(np.max(np.abs(E_mix - E_ref)))
    print("mixed(%d) iteration errors: W %.2e  Sigma %.2e  Ec %.2e Ha  QP %.3f meV"
          % (prec, e_w, e_s, e_ec, e_qp / HA_PER_MEV))
    assert e_w < TOL_W and e_s < TOL_S
    assert e_ec < TOL_EC
    assert e_qp < TOL_QP


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("kgrid,table,no,nmu,tol_w", [
    ((3, 3, 3), "1e2.1e-15", 9, 37, TOL_W),
    ((4, 4, 4), "1e5.1e-06", 7, 24, 2e-4),
    ((2, 3, 1), "1e2.1e-15", 6, 20, TOL_W),      # generic-grid transform kernel
    ((2, 2, 2), "1e2.1e-15", 10, 130, TOL_W),    # N_mu past two tiles: partial tiles, K chunks
])
def test_mixed_iteration_shapes(orc, torch, kgrid, table, no, nmu, tol_w, prec):
    s, ia, tab, G = system(orc, kgrid, no, nmu, table, sid=6)
    ref = orc.gw_iteration(tab, G, s.X, s.V, s.kgrid, with_ec=True)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    ctx.set_flag(7, prec)
    S, W, ec = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
    e_w, e_s, e_ec = relmax(to_np(W), ref["W_qt"]), relmax(to_np(S), ref["Sigma"]), abs(ec - ref["Ec"])
    print("mixed(%d) errors %s nmu=%d: W %.2e  Sigma %.2e  Ec %.2e Ha" % (prec, table, nmu, e_w, e_s, e_ec))
    assert e_w < tol_w and e_s < TOL_S and e_ec < TOL_EC
