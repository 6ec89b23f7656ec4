"""Batched form: independent systems of uneven size iterated together, captured into one
HIP graph and replayed (thcgw_batch_*).

Every replay is checked against the same systems run one by one through
thcgw_gw_iteration (1e-13 relative: the graph replays the same kernels on the same
inputs), with eager work and synchronisations between the replays, and one system of the
plain FP64 batch against the oracle (1e-10).
"""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcBatch, ThcGW, make_system

pytestmark = pytest.mark.gpu

SYSTEMS = [((2, 2, 2), 10, 80), ((3, 3, 3), 8, 64), ((2, 2, 2), 14, 112), ((1, 1, 1), 6, 48)]
TABLE = "1e2.1e-15"


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


def build(orc, torch, precision, trev=False, dyson=0):
    ia = IAFT.from_table(1000.0, TABLE)
    tab = orc.load_ir(1000.0, None, table=TABLE)
    out = []
    for sid, (kgrid, no, nmu) in enumerate(SYSTEMS):
        s = make_system(kgrid, no, nmu, system_id=40 + sid, trev=trev)
        ctx = ThcGW(kgrid, no, nmu, ia)
        ctx.set_flag(6, dyson)
        ctx.set_flag(10, int(dyson == 2))  # the resolvent batch also projects G as Hermitian
        ctx.set_flag(7, precision)
        ctx.set_flag(9, int(trev))
        G = torch.from_numpy(orc.update_G(tab, s.C, s.E)).cuda()
        X = torch.from_numpy(s.X).cuda()
        V = torch.from_numpy(s.V).cuda()
        Sig = torch.empty((ia.nt, 1, s.nk, no, no), dtype=torch.complex128, device="cuda")
        W = torch.empty((s.nk, ia.nt_half, nmu, nmu), dtype=torch.complex128, device="cuda")
        out.append(dict(s=s, ctx=ctx, G=G, X=X, V=V, Sig=Sig, W=W, tab=tab))
    return out


def eager(sys_):
    S, W, ec = sys_["ctx"].gw_iteration(sys_["G"], sys_["X"], sys_["V"], with_ec=True, hermitize=True)
    return S.cpu().numpy(), W.cpu().numpy(), ec


@pytest.mark.parametrize("precision,trev,dyson,lanes", [(0, False, 0, 1), (1, False, 0, 1), (0, True, 0, 1),
                                                       (1, False, 2, 1), (0, False, 2, 1), (1, False, 2, 2),
                                                       (0, False, 0, 3)])
def test_batch_graph_replays_match_eager(orc, torch, precision, trev, dyson, lanes):
    """With trev, every system is time-reversal symmetric and solves the irreducible q only
    (flag 9), inside the captured graph too; dyson = 2 runs the resolvent Dyson form; with
    lanes > 1 the systems are spread over that many streams, parallel branches of the graph
    (each lane with its own scratch)."""
    print("case: batch precision=%d trev=%d dyson=%d lanes=%d systems=%s" % (precision, trev, dyson, lanes, SYSTEMS),
          flush=True)
    systems = build(orc, torch, precision, trev, dyson)
    b = ThcBatch([x["ctx"] for x in systems], use_graph=True, tune=(precision != 0), lanes=lanes)
    args = [[x[k] for x in systems] for k in ("G", "X", "V", "Sig", "W")]
    try:
        ec0 = b.gw_iteration(*args)  # eager pass (and tile tuning)
        refs = [eager(x) for x in systems]  # one by one, with the tuned tiles
        for rep in range(3):
            ec = b.gw_iteration(*args)  # captured on the first of these, replayed after
            torch.cuda.synchronize()
            for x, (S, W, e), e_b in zip(systems, refs, ec):
                assert relmax(x["Sig"].cpu().numpy(), S) < 1e-13
                assert relmax(x["W"].cpu().numpy(), W) < 1e-13
                assert abs(e_b - e) < 1e-12 * abs(e)
            # eager work between replays: one system alone, then a host round trip
            eager(systems[rep % len(systems)])
            torch.cuda.synchronize()
        for e_a, (_, _, e) in zip(ec0, refs):
            assert abs(e_a - e) < 1e-12 * abs(e)
        if precision == 0:  # one system against the oracle (resolvent form: 1e-10 + cond(V) eps)
            x = systems[1]
            ref = orc.gw_iteration(x["tab"], x["G"].cpu().numpy(), x["s"].X, x["s"].V, x["s"].kgrid, with_ec=True)
            tol = 1e-10 if dyson != 2 else 1e-10 + max(np.linalg.cond(v) for v in x["s"].V) * np.finfo(float).eps
            assert relmax(x["Sig"].cpu().numpy(), ref["Sigma"]) < tol
            assert relmax(x["W"].cpu().numpy(), ref["W_qt"]) < tol
    finally:
        b.close()


def test_batch_graph_follows_gemm_mode(orc, torch):
    """The process-wide GEMM settings are part of the captured graph's key: after flag 12
    (complex products by three real products) changes, the next batch call captures again
    and matches the systems run one by one with the new setting (1e-13), not the graph of
    the old one."""
    print("case: batch graph and flag 12 systems=%s" % (SYSTEMS,), flush=True)
    systems = build(orc, torch, 1, False, 2)
    b = ThcBatch([x["ctx"] for x in systems], use_graph=True, tune=False)
    args = [[x[k] for x in systems] for k in ("G", "X", "V", "Sig", "W")]
    try:
        for mode in (0, 13, 0):
            systems[0]["ctx"].set_flag(12, mode)
            b.gw_iteration(*args)  # eager pass (first call) or capture
            refs = [eager(x) for x in systems]
            for _ in range(2):
                b.gw_iteration(*args)
                torch.cuda.synchronize()
                for x, (S, W, e) in zip(systems, refs):
                    assert relmax(x["Sig"].cpu().numpy(), S) < 1e-13, mode
                    assert relmax(x["W"].cpu().numpy(), W) < 1e-13, mode
    finally:
        systems[0]["ctx"].set_flag(12, 0)
        b.close()
