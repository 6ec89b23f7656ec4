"""The BLAS-backed CPU port used as the timed CPU baseline agrees with the oracle."""
import numpy as np

from coqui_amd.system import make_system
from oracle import blas_port, oracle


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def setup():
    kgrid = (2, 2, 2)
    s = make_system(kgrid, 6, 20, system_id=3)
    tab = oracle.load_ir(1000.0, None, table="1e2.1e-15")
    G = oracle.update_G(tab, s.C, s.E)
    return kgrid, s, tab, G


def test_pi_and_sigma_pair():
    kgrid, s, tab, G = setup()
    Gp = np.ascontiguousarray(G[[0, tab["nt"] - 1]])
    ref = oracle.pi_rspace(Gp, s.X, kgrid)[0]
    assert rel(blas_port.pi_pair(Gp, s.X, kgrid), ref) < 1e-12
    W = np.random.default_rng(2).standard_normal(ref.shape) + 1j * np.random.default_rng(3).standard_normal(ref.shape)
    Sref = oracle.sigma_rspace(Gp, s.X, np.ascontiguousarray(W[:, None]), kgrid)[:, 0]
    assert rel(blas_port.sigma_pair(Gp, s.X, W, kgrid), Sref) < 1e-12


def test_dyson_terms_and_transforms():
    kgrid, s, tab, G = setup()
    Pit = oracle.pi_rspace(G, s.X, kgrid)
    Piw = oracle.tau_to_w_phsym(tab, Pit)
    Tp, Tm = blas_port.fold_tables(tab)
    assert rel(blas_port.tau_to_w(Tp, Pit), Piw) < 1e-12
    W = oracle.dyson(Piw, s.V)
    assert rel(blas_port.dyson(Piw[:3], s.V), W[:3]) < 1e-11
    assert rel(blas_port.w_to_tau(Tm, W), oracle.w_to_tau_phsym(tab, W)) < 1e-12
    tr, ld = blas_port.rpa_terms(Piw[:2], s.V)
    for w in range(2):
        for q in range(s.nk):
            A = Piw[w, q] @ s.V[q]
            assert abs(tr[w, q] - np.trace(A)) < 1e-12 * abs(np.trace(A))
            assert abs(ld[w, q] - np.linalg.slogdet(np.eye(A.shape[0]) - A)[1]) < 1e-10
