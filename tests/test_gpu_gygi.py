"""Long-wavelength (Gygi) treatment on the GPU against the FP64 oracle (1e-10):
thcgw_eps_inv_head (g0_div_utils.hpp:55-111,180-322) in the three modes and
thcgw_sigma_div_correction (thc_gw.icc:446-526); ignore_g0 and a single k-point leave Sigma
untouched."""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, make_system

pytestmark = pytest.mark.gpu

TABLE = "1e2.1e-15"


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def qcart(kgrid, L=10.0):
    out = []
    for i in range(kgrid[0]):
        for j in range(kgrid[1]):
            for l in range(kgrid[2]):
                f = np.array([i / kgrid[0], j / kgrid[1], l / kgrid[2]])
                out.append(2 * np.pi * (f - np.round(f)) / L)
    return np.array(out)


@pytest.mark.parametrize("kgrid,no,nmu", [((4, 4, 4), 8, 40), ((3, 3, 1), 6, 33)])
def test_gygi_matches_oracle(orc, kgrid, no, nmu):
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    print("case: gygi kgrid=%s no=%d nmu=%d" % (kgrid, no, nmu), flush=True)
    s = make_system(kgrid, no, nmu, system_id=91)
    tab = orc.load_ir(1000.0, None, table=TABLE)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    G = ctx.update_G(s.C, s.E, 0.0)
    S, W, _ = ctx.gw_iteration(G, s.X, s.V, with_ec=False)
    rng = np.random.default_rng(12)
    cb = rng.standard_normal((s.nk, nmu)) + 1j * rng.standard_normal((s.nk, nmu))
    ch = rng.standard_normal(nmu) + 1j * rng.standard_normal(nmu)
    q = qcart(kgrid)
    Wh = W.cpu().numpy()
    Gh = G.cpu().numpy()
    for mode in (1, 2, 3):
        tq, t0 = ctx.eps_inv_head(W, cb, q, 750.0, mode)
        rtq, rt0 = orc.eps_inv_head(tab, Wh, cb, q, 750.0, mode)
        print("mode %d: eps(t,q) %.1e  head %.1e" % (mode, relmax(tq, rtq), relmax(t0, rt0)))
        assert relmax(tq, rtq) < 1e-10 and relmax(t0, rt0) < 1e-10
        S2 = S.clone()
        ctx.sigma_div_correction(S2, G, s.X, ch, 2.3, t0, mode)
        ref = orc.sigma_div_correction(S.cpu().numpy(), Gh, s.X, ch, 2.3, rt0, mode)
        print("mode %d: Sigma %.1e (correction %.2e of Sigma)" % (
            mode, relmax(S2.cpu().numpy(), ref), relmax(S2.cpu().numpy(), S.cpu().numpy())))
        assert relmax(S2.cpu().numpy(), ref) < 1e-10
    S3 = S.clone()
    ctx.sigma_div_correction(S3, G, s.X, ch, 2.3, t0, "ignore_g0")
    assert torch.equal(S3, S)


def test_gygi_single_kpoint_is_noop(orc):
    import torch

    kgrid, no, nmu = (1, 1, 1), 6, 20
    s = make_system(kgrid, no, nmu, system_id=92)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    G = ctx.update_G(s.C, s.E, 0.0)
    S, W, _ = ctx.gw_iteration(G, s.X, s.V, with_ec=False)
    cb = np.ones((1, nmu), complex)
    tq, t0 = ctx.eps_inv_head(W, cb, np.zeros((1, 3)), 100.0, "gygi_extrplt")
    assert np.all(tq == 0)  # |q| = 0 only
    S2 = S.clone()
    ctx.sigma_div_correction(S2, G, s.X, np.ones(nmu, complex), 2.0, t0, "gygi_extrplt")
    assert torch.equal(S2, S)
