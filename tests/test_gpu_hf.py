"""THC Hartree-Fock potential (thcgw_hf) against the FP64 oracle (thc_hf.icc:18-300).

Tolerance: 1e-10 relative to the largest element of the oracle result, as for the plain
FP64 GW path.
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


def density_matrix(s, beta=1000.0):
    f = 1.0 / (1.0 + np.exp(np.clip(beta * s.E, -700, 700)))
    return np.einsum("kia,ka,kja->kij", s.C, f, s.C.conj())


@pytest.mark.parametrize("kgrid,no,nmu", [
    ((2, 2, 2), 26, 200),     # N_k = 8, N_orb = 26, N_mu = 200
    ((4, 4, 4), 12, 70),      # lane-split N_k = 64 transform
    ((3, 3, 3), 9, 37),
    ((1, 1, 1), 7, 50),       # Gamma only
    ((2, 3, 1), 6, 20),       # generic-grid transform kernel
])
def test_hf_matches_oracle(orc, torch, kgrid, no, nmu):
    print("case: hf kgrid=%s no=%d nmu=%d" % (kgrid, no, nmu), flush=True)
    s = make_system(kgrid, no, nmu, system_id=8)
    Dm = density_matrix(s)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, "1e2.1e-15"))
    for hartree, exchange in ((True, True), (True, False), (False, True)):
        ref = orc.thc_hf(Dm, s.X, s.V, kgrid, hartree=hartree, exchange=exchange)
        F = ctx.hf(Dm[None], s.X, s.V, hartree=hartree, exchange=exchange)
        err = relmax(F.cpu().numpy()[0], ref)
        print("hf hartree=%d exchange=%d err %.2e" % (hartree, exchange, err))
        assert err < TOL
