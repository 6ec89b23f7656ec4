"""q-point symmetry map (thcgw_set_qsym): Dyson on the irreducible q, the rest filled by
phased permutations of the THC points.  Checked against the same context without the map
(1e-12) and the FP64 oracle (1e-10), the Ec terms, time reversal expressed as a map against
flag 9, and the refusal of input that does not follow the map."""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, ThcgwError, make_system
from coqui_amd.system import minus_k

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


@pytest.mark.parametrize("kgrid,no,nmu", [((4, 4, 1), 8, 40), ((2, 2, 2), 6, 48), ((4, 4, 4), 5, 24)])
def test_qsym_matches_full_and_oracle(orc, torch, kgrid, no, nmu):
    print("case: qsym c4 kgrid=%s no=%d nmu=%d" % (kgrid, no, nmu), flush=True)
    s = make_system(kgrid, no, nmu, system_id=101, c4=True)
    src, perm = s.qsym
    tab = orc.load_ir(1000.0, None, table=TABLE)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    G = ctx.update_G(s.C, s.E, 0.0)
    ref = orc.gw_iteration(tab, orc.update_G(tab, s.C, s.E), s.X, s.V, kgrid, with_ec=True)
    for f6 in (0, 1, 2):
        ctx.set_flag(6, f6)
        ctx.set_qsym()
        Sf, Wf, ef = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
        ctx.set_qsym(src, perm)
        St, Wt, et = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
        Sf, Wf, St, Wt = (a.cpu().numpy() for a in (Sf, Wf, St, Wt))
        print("form %d: nir %d of %d, W %.1e Sigma %.1e Ec %.1e" % (
            f6, len(set(src.tolist())), s.nk, relmax(Wt, Wf), relmax(St, Sf), abs(et - ef)))
        # the resolvent form rounds with cond(V) (its bound: 1e-10 + cond(V) eps)
        t = 1e-12 if f6 < 2 else 1e-10
        to = 1e-10 if f6 < 2 else 1e-10 + max(np.linalg.cond(v) for v in s.V) * np.finfo(float).eps
        assert relmax(Wt, Wf) < t and relmax(St, Sf) < t and abs(et - ef) < t * abs(ef)
        assert relmax(Wt, ref["W_qt"]) < to and relmax(St, ref["Sigma"]) < to
        assert abs(et - ref["Ec"]) < 1e-10 * abs(ref["Ec"])
    ctx.set_flag(6, 1)
    ctx.set_flag(7, 1)  # mixed precision with the map
    St, Wt, et = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
    assert relmax(Wt.cpu().numpy(), ref["W_qt"]) < 5e-5 and abs(et - ref["Ec"]) < 1e-6
    ctx.set_flag(7, 0)
    ctx.set_qsym()


def test_qsym_phases_and_terms(orc, torch):
    """Phases: Pi(iW) and V of every mapped q taken to D_q A D_q^H (D_q diagonal unit
    phases) follow the map with those phases; thcgw_dyson_w and thcgw_rpa_ec with the map
    agree with the unmapped solve of the same input (W, Tr, ln det to 1e-12)."""
    kgrid, no, nmu = (4, 4, 1), 6, 32
    s = make_system(kgrid, no, nmu, system_id=102, c4=True)
    src, perm = s.qsym
    rng = np.random.default_rng(3)
    ph = np.exp(2j * np.pi * rng.random((s.nk, nmu)))
    ph[src == np.arange(s.nk)] = 1.0
    D = lambda A: A * ph[:, :, None] * ph.conj()[:, None, :]  # per q: D_q A D_q^H
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    G = ctx.update_G(s.C, s.E, 0.0)
    Piw = ctx.tau_to_w(ctx.pi(G, s.X)).cpu().numpy()
    Piw = torch.from_numpy(np.ascontiguousarray(D(Piw))).cuda()
    V = np.ascontiguousarray(D(s.V))
    Wf = Piw.clone()
    tr_f, ld_f = ctx.dyson_w(Wf, V, with_terms=True)
    ec_f = ctx.rpa_ec(Piw, V)
    ctx.set_qsym(src, perm, ph)
    Wt = Piw.clone()
    tr_t, ld_t = ctx.dyson_w(Wt, V, with_terms=True)
    ec_t = ctx.rpa_ec(Piw, V)
    ctx.set_qsym()
    assert relmax(Wt.cpu().numpy(), Wf.cpu().numpy()) < 1e-12
    assert np.max(np.abs(tr_t - tr_f)) < 1e-12 * np.max(np.abs(tr_f))
    assert np.max(np.abs(ld_t - ld_f)) < 1e-12 * np.max(np.abs(ld_f))
    assert abs(ec_t - ec_f) < 1e-12 * abs(ec_f)


def test_qsym_time_reversal_equals_flag9(orc, torch):
    kgrid, no, nmu = (3, 3, 1), 6, 30
    s = make_system(kgrid, no, nmu, system_id=103, trev=True)
    mk = minus_k(kgrid)
    src = np.minimum(np.arange(s.nk), mk)
    conj = (src != np.arange(s.nk)).astype(np.int32)
    perm = np.tile(np.arange(nmu, dtype=np.int32), (s.nk, 1))
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    G = ctx.update_G(s.C, s.E, 0.0)
    ctx.set_flag(9, 1)
    S9, W9, e9 = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
    ctx.set_qsym(src, perm, None, conj)
    Sq, Wq, eq = ctx.gw_iteration(G, s.X, s.V, with_ec=True)
    assert torch.equal(W9, Wq) and abs(e9 - eq) < 1e-13 * abs(e9)
    assert relmax(Sq.cpu().numpy(), S9.cpu().numpy()) < 1e-13


def test_qsym_refuses_broken_map(torch):
    kgrid, no, nmu = (4, 4, 1), 6, 24
    s = make_system(kgrid, no, nmu, system_id=104, c4=True)
    src, perm = s.qsym
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    bad = perm.copy()
    q = int(np.nonzero(src != np.arange(s.nk))[0][0])
    bad[q, [0, 1]] = bad[q, [1, 0]]
    ctx.set_qsym(src, bad)
    G = ctx.update_G(s.C, s.E, 0.0)
    with pytest.raises(ThcgwError, match="does not follow the map"):
        ctx.gw_iteration(G, s.X, s.V, with_ec=False)
    with pytest.raises(ThcgwError, match="irreducible"):
        ctx.set_qsym(np.roll(src, 1), perm)
