"""Dyson equation for G, electron number and chemical-potential search on the GPU against
the FP64 oracle (simple_dyson.cpp:11-172, scf_common.cpp:10-42,147-202).

Tolerances: G and Dm 1e-10 relative to the largest oracle element; nelec 1e-10 relative;
the chemical potential reproduces the target electron number within mu_tol = 1e-9 and
agrees with the oracle's to 1e-6 Ha (both bisect the same bracket; their N(mu) differ by
rounding only).
"""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, make_system

pytestmark = pytest.mark.gpu

TOL = 1e-10


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


def pole_sigma(tab, s, ep=0.3, scale=0.02, seed=5):
    """A causal self-energy: one pole at ep per k with a Hermitian positive weight."""
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((s.nk, s.no, s.no)) + 1j * rng.standard_normal((s.nk, s.no, s.no))
    Wt = scale * B @ np.conj(np.swapaxes(B, 1, 2)) / s.no
    g = -np.exp(-ep * tab["tau"]) / (1.0 + np.exp(-ep * tab["beta"]))
    return np.ascontiguousarray((g[:, None, None, None] * Wt[None])[:, None])


@pytest.mark.parametrize("kgrid,no,table", [((2, 2, 1), 7, "1e4.1e-15"), ((2, 2, 2), 26, "1e3.1e-15")])
def test_dyson_g_matches_oracle(orc, torch, kgrid, no, table):
    print("case: dyson_g kgrid=%s no=%d table=%s" % (kgrid, no, table), flush=True)
    s = make_system(kgrid, no, 4, system_id=34)
    tab = orc.load_ir(1000.0, None, table=table)
    ctx = ThcGW(kgrid, no, 4, IAFT.from_table(1000.0, table))
    Sig = pole_sigma(tab, s)
    H = s.H[None]
    for mu in (0.0, 0.07):
        G, Dm = ctx.dyson_g(H, Sig, mu)
        Gr, Dmr = orc.dyson_g(tab, s.H, Sig, mu)
        assert relmax(G.cpu().numpy(), Gr) < TOL
        assert relmax(Dm.cpu().numpy()[0], Dmr) < TOL
        n_gpu = ctx.nelec(H, Sig, mu)
        n_orc = orc.nelec(tab, s.H, Sig, mu)
        assert abs(n_gpu - n_orc) < TOL * max(1.0, abs(n_orc))


def test_dyson_g_free_particle(orc, torch):
    """Sigma = 0, F = H(k): G(tau) equals update_G's closed form (lambda = 1e4 covers |e| <= 1)."""
    s = make_system((2, 1, 1), 9, 4, system_id=35)
    tab = orc.load_ir(1000.0, None, table="1e4.1e-15")
    ctx = ThcGW((2, 1, 1), 9, 4, IAFT.from_table(1000.0, "1e4.1e-15"))
    Sig = np.zeros((tab["nt"], 1, s.nk, s.no, s.no), complex)
    G, Dm = ctx.dyson_g(s.H[None], Sig, 0.0)
    assert relmax(G.cpu().numpy(), orc.update_G(tab, s.C, s.E, 0.0)) < 1e-9


def test_update_mu(orc, torch):
    kgrid, no = (2, 2, 1), 8
    s = make_system(kgrid, no, 4, system_id=36)
    tab = orc.load_ir(1000.0, None, table="1e4.1e-15")
    ctx = ThcGW(kgrid, no, 4, IAFT.from_table(1000.0, "1e4.1e-15"))
    Sig = pole_sigma(tab, s, scale=0.01)
    target = 7.3  # electrons per cell, between integer fillings
    mu_g, n_g = ctx.update_mu(s.H[None], Sig, target, 0.0, mu_tol=1e-9)
    mu_o, n_o = orc.update_mu(tab, s.H, Sig, target, 0.0, mu_tol=1e-9)
    print("update_mu gpu %.12f (N %.12f) oracle %.12f (N %.12f)" % (mu_g, n_g, mu_o, n_o))
    assert abs(n_g - target) < 1e-9
    assert abs(mu_g - mu_o) < 1e-6
