"""Brillouin-zone symmetry on the GPU (thcgw_set_bz, thcgw_pi_ibz, thcgw_sigma_ibz,
thcgw_gw_iteration_ibz).

On symmetric systems (a fourfold rotation whose high-symmetry points carry non-trivial
orbital representations, and time reversal):

* chi0 on the irreducible q and Sigma on the irreducible k match the wedge oracle
  (oracle/bz_oracle.py, itself pinned against the full-grid oracle) to 1e-10;
* the whole iteration on the wedge matches the full-grid iteration of the same system at
  the irreducible points: Sigma, W and Ec to 1e-10 (LU and resolvent Dyson forms);
* malformed tables and unsupported settings are refused before any kernel runs.
"""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, ThcgwError
from coqui_amd.system import make_system_bz
from oracle import bz_oracle

pytestmark = pytest.mark.gpu

TABLE = "1e2.1e-15"
CASES = [("c4", (3, 3, 1), 8, 48), ("c4", (4, 4, 1), 6, 40), ("c4", (4, 4, 4), 6, 32), ("trev", (3, 2, 2), 8, 48),
         ("trev", (2, 2, 3), 6, 32)]


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


def setup(torch, kind, kgrid, no, nmu):
    s, bz = make_system_bz(kgrid, no, nmu, system_id=71, kind=kind)
    ia = IAFT.from_table(1000.0, TABLE)
    ctx = ThcGW(kgrid, no, nmu, ia)
    ctx.set_bz(bz)
    order = bz["order"]
    G = ctx.update_G(s.C, s.E, s.mu)  # full grid, C order
    Gi = G[:, :, torch.from_numpy(order[:bz["nk_ibz"]]).cuda()].contiguous()
    Xl = torch.from_numpy(np.ascontiguousarray(s.X[order])).cuda()
    return s, bz, ia, ctx, G, Gi, Xl


@pytest.mark.parametrize("kind,kgrid,no,nmu", CASES)
def test_pi_sigma_ibz_vs_oracle(torch, kind, kgrid, no, nmu):
    print("case: bz pi/sigma %s kgrid=%s no=%d nmu=%d" % (kind, kgrid, no, nmu), flush=True)
    s, bz, ia, ctx, G, Gi, Xl = setup(torch, kind, kgrid, no, nmu)
    try:
        Gi_h, Xl_h = Gi.cpu().numpy(), Xl.cpu().numpy()
        Pi_ref = bz_oracle.pi_ibz(Gi_h, Xl_h, bz, kgrid)
        rng = np.random.default_rng(5)
        ct = rng.uniform(0.5, 1.5, ia.nt_half)
        Wi = np.ascontiguousarray(s.V[bz["order"][:bz["nq_ibz"]]][:, None] * ct[None, :, None, None])
        S_ref = bz_oracle.sigma_ibz(Gi_h, Xl_h, Wi, bz)
        for herm in (0, 1):  # flag 10: Hermitian G projections and Hermitian Sigma_PQ halves
            ctx.set_flag(10, herm)
            P = ctx.pi_ibz(Gi, Xl)
            assert relmax(P.cpu().numpy(), Pi_ref) < 1e-10
            S = ctx.sigma_ibz(Gi, torch.from_numpy(Wi).cuda(), Xl, hermitize=False)
            assert relmax(S.cpu().numpy(), S_ref) < 1e-10
        # against the full-grid GPU path at the irreducible q
        ctx.set_flag(10, 0)
        Pf = ctx.pi(G, torch.from_numpy(s.X).cuda())
        ref = Pf[:, torch.from_numpy(bz["order"][:bz["nq_ibz"]]).cuda()]
        assert relmax(P.cpu().numpy(), ref.cpu().numpy()) < 1e-10
    finally:
        ctx.close()


@pytest.mark.parametrize("kind,kgrid,no,nmu,dyson", [c + (0,) for c in CASES] + [CASES[0] + (2,), CASES[3] + (2,)])
def test_iteration_ibz_vs_full_grid(torch, kind, kgrid, no, nmu, dyson):
    """dyson = 2 runs the resolvent form with flag 10 (Hermitian G), as the benchmark does."""
    print("case: bz iteration %s kgrid=%s no=%d nmu=%d dyson=%d" % (kind, kgrid, no, nmu, dyson), flush=True)
    s, bz, ia, ctx, G, Gi, Xl = setup(torch, kind, kgrid, no, nmu)
    try:
        ctx.set_flag(6, dyson)
        ctx.set_flag(10, int(dyson == 2))
        V = torch.from_numpy(s.V).cuda()
        Sf, Wf, ecf = ctx.gw_iteration(G, torch.from_numpy(s.X).cuda(), V)
        order = torch.from_numpy(bz["order"]).cuda()
        Vi = V[order[:bz["nq_ibz"]]].contiguous()
        Si, Wi, eci = ctx.gw_iteration_ibz(Gi, Xl, Vi)
        assert relmax(Si.cpu().numpy(), Sf[:, :, order[:bz["nk_ibz"]]].cpu().numpy()) < 1e-10
        assert relmax(Wi.cpu().numpy(), Wf[order[:bz["nq_ibz"]]].cpu().numpy()) < 1e-10
        assert abs(eci - ecf) < 1e-10 * abs(ecf)
        # a second call reuses the workspaces
        Si2, _, eci2 = ctx.gw_iteration_ibz(Gi, Xl, Vi)
        assert relmax(Si2.cpu().numpy(), Si.cpu().numpy()) < 1e-14 and eci2 == eci
    finally:
        ctx.close()


def test_bz_tables_refused(torch):
    print("case: bz refused tables", flush=True)
    s, bz, ia, ctx, G, Gi, Xl = setup(torch, "trev", (3, 2, 2), 6, 24)
    try:
        bad = dict(bz, kp_to_ibz=np.where(np.arange(12) == 11, 99, bz["kp_to_ibz"]))
        with pytest.raises(ThcgwError, match="kp_to_ibz"):
            ctx.set_bz(bad)
        with pytest.raises(ThcgwError, match="no Brillouin-zone symmetry"):
            ctx.pi_ibz(Gi, Xl)  # the failed set cleared the symmetry
        ks = bz["ks_to_k"].copy()
        ks[0, 0] = 11  # a time-reversal point
        with pytest.raises(ThcgwError, match="k S"):
            ctx.set_bz(dict(bz, ks_to_k=ks))
        ctx.set_bz(bz)
        ctx.set_flag(7, 1)
        with pytest.raises(ThcgwError, match="FP64"):
            ctx.pi_ibz(Gi, Xl)
    finally:
        ctx.close()
