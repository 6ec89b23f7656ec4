"""Time-reversal symmetric input (flag 9): the Dyson step solves the irreducible q only and
fills W(-q) = conj W(q).  Checked against the same context without the flag (1e-12) and the
FP64 oracle (1e-10) for the LU and Hermitian forms (the resolvent form to 1e-10 against the
full path and to 1e-10 + cond(V) eps against the oracle), against the oracle at the mixed form's
tolerances, for the stage entry points and Ec; input without the symmetry is refused."""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, ThcgwError, make_system

pytestmark = pytest.mark.gpu

TABLE = "1e2.1e-15"


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run(ctx, s, G, flag6, flag7, trev):
    ctx.set_flag(6, flag6)
    ctx.set_flag(7, flag7)
    ctx.set_flag(9, int(trev))
    S, W, ec = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
    return S.cpu().numpy(), W.cpu().numpy(), ec


@pytest.mark.parametrize("kgrid,no,nmu", [((2, 2, 2), 10, 48), ((3, 3, 1), 8, 40), ((4, 4, 4), 6, 24)])
def test_trev_matches_full_and_oracle(orc, torch, kgrid, no, nmu):
    print("case: trev kgrid=%s no=%d nmu=%d" % (kgrid, no, nmu), flush=True)
    s = make_system(kgrid, no, nmu, system_id=81, trev=True)
    tab = orc.load_ir(1000.0, None, table=TABLE)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    G = ctx.update_G(s.C, s.E, 0.0)
    ref = orc.gw_iteration(tab, orc.update_G(tab, s.C, s.E), s.X, s.V, kgrid, with_ec=True)
    for f6 in (0, 1, 2):
        Sf, Wf, ef = run(ctx, s, G, f6, 0, False)
        St, Wt, et = run(ctx, s, G, f6, 0, True)
        print("form %d: W %.1e Sigma %.1e Ec %.1e" % (f6, relmax(Wt, Wf), relmax(St, Sf), abs(et - ef)))
        # the resolvent form rounds with cond(V) (its bound: 1e-10 + cond(V) eps)
        t = 1e-12 if f6 < 2 else 1e-10
        to = 1e-10 if f6 < 2 else 1e-10 + max(np.linalg.cond(v) for v in s.V) * np.finfo(float).eps
        assert relmax(Wt, Wf) < t and relmax(St, Sf) < t and abs(et - ef) < t * abs(ef)
        assert relmax(Wt, ref["W_qt"]) < to and relmax(St, ref["Sigma"]) < to
        assert abs(et - ref["Ec"]) < 1e-10 * abs(ref["Ec"])
    St, Wt, et = run(ctx, s, G, 1, 1, True)  # mixed precision
    assert relmax(Wt, ref["W_qt"]) < 5e-5 and relmax(St, ref["Sigma"]) < 5e-5 and abs(et - ref["Ec"]) < 1e-6
    ctx.set_flag(7, 0)


def test_trev_stage_entry_points(orc, torch):
    """thcgw_dyson_w (with its Ec terms) and thcgw_rpa_ec under flag 9."""
    kgrid, no, nmu = (2, 2, 2), 8, 40
    s = make_system(kgrid, no, nmu, system_id=82, trev=True)
    tab = orc.load_ir(1000.0, None, table=TABLE)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    G = ctx.update_G(s.C, s.E, 0.0)
    Piw = ctx.tau_to_w(ctx.pi(G, s.X))
    P0 = Piw.clone()
    ec_full = ctx.rpa_ec(Piw, s.V)
    tr_f, ld_f = ctx.dyson_w(P0.clone(), s.V, with_terms=True)
    ctx.set_flag(9, 1)
    ec_tr = ctx.rpa_ec(Piw, s.V)
    W = P0.clone()
    tr_t, ld_t = ctx.dyson_w(W, s.V, with_terms=True)
    assert torch.equal(Piw, P0)  # rpa_ec leaves Pi untouched
    assert abs(ec_tr - ec_full) < 1e-12 * abs(ec_full)
    assert np.max(np.abs(tr_t - tr_f)) < 1e-12 * np.max(np.abs(tr_f)) and np.max(np.abs(ld_t - ld_f)) < 1e-12 * np.max(np.abs(ld_f))
    ref = orc.dyson(P0.cpu().numpy(), s.V)
    assert relmax(W.cpu().numpy(), ref) < 1e-10


def test_trev_refuses_asymmetric_input(torch):
    kgrid, no, nmu = (2, 2, 1), 6, 24
    s = make_system(kgrid, no, nmu, system_id=83)  # no symmetry
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    ctx.set_flag(9, 1)
    G = ctx.update_G(s.C, s.E, 0.0)
    with pytest.raises(ThcgwError, match="not symmetric"):
        ctx.gw_iteration(G, s.X, s.V, with_ec=False)
