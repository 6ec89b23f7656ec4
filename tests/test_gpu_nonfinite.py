"""Non-finite input and overflow in the Dyson step end in an error code, not in a fault.

* NaN / infinite inputs are refused by the entry points' input guard before any
  factorisation (THCGW_ERR_NONFINITE = 6).
* Finite input whose product I - V Pi overflows reaches the LU factorisation with
  infinite entries: the panel then pivots on an in-range row (its candidate index starts
  at the first row of the column, as LAPACK's izamax does), flags the non-finite pivot,
  and the call returns THCGW_ERR_SINGULAR = 3.
* The context keeps working afterwards (a finite call matches the oracle).
"""
import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW, ThcgwError, make_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


@pytest.fixture(scope="module")
def small(orc, torch):
    kgrid, no, nmu, table = (2, 1, 1), 6, 40, "1e2.1e-15"
    s = make_system(kgrid, no, nmu, system_id=5)
    ia = IAFT.from_table(1000.0, table)
    tab = orc.load_ir(1000.0, None, table=table)
    return s, ia, tab


def _pi_w(s, ia, torch, scale=1.0):
    rng = np.random.default_rng(11)
    B = rng.standard_normal((ia.nw_half, s.nk, s.nmu, s.nmu)) + 1j * rng.standard_normal((ia.nw_half, s.nk, s.nmu, s.nmu))
    Pi = -(B @ np.conj(np.swapaxes(B, -1, -2))) / s.nmu * 0.05 * scale
    return torch.from_numpy(np.ascontiguousarray(Pi)).cuda()


@pytest.mark.parametrize("algo", [0, 1, 2])
def test_nan_input_is_refused(small, torch, algo):
    s, ia, tab = small
    print("case: NaN input, Dyson algorithm %d" % algo, flush=True)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    ctx.set_flag(6, algo)
    V = torch.from_numpy(s.V).cuda()
    Pi = _pi_w(s, ia, torch)
    Pi[1, 0, 3, 7] = complex(float("nan"), 0.0)
    with pytest.raises(ThcgwError) as e:
        ctx.dyson_w(Pi, V)
    assert e.value.code == 6 and "Pi_wqPQ" in str(e.value)
    Vb = V.clone()
    Vb[0, 2, 2] = complex(float("inf"), 0.0)
    G = ctx.update_G(s.C, s.E, 0.0)
    with pytest.raises(ThcgwError) as e:
        ctx.gw_iteration(G, s.X, Vb)
    assert e.value.code == 6 and "V_qPQ" in str(e.value)
    ctx.close()


def test_overflow_in_lu_dyson_is_singular(small, orc, torch):
    """Finite V and Pi = -c I with c = 1.6e308: the diagonal of I - V Pi overflows to
    infinity in the product, so the LU panel meets an infinite pivot in its first column."""
    s, ia, tab = small
    print("case: overflowing I - V Pi, LU Dyson form", flush=True)
    ctx = ThcGW(s.kgrid, s.no, s.nmu, ia)
    ctx.set_flag(6, 0)
    V = torch.from_numpy(s.V).cuda()
    eye = torch.eye(s.nmu, dtype=torch.complex128, device="cuda")
    Pi = (-1.6e308 * eye).expand(ia.nw_half, s.nk, s.nmu, s.nmu).contiguous()
    V2 = (2.0 * V).contiguous()
    assert bool(torch.isfinite(Pi).all()) and bool(torch.isfinite(V2).all())
    assert float(V2.diagonal(dim1=1, dim2=2).real.min()) * 1.6e308 == float("inf")
    with pytest.raises(ThcgwError) as e:
        ctx.dyson_w(Pi, V2)
    assert e.value.code == 3
    # the context still works: a finite call matches the oracle
    Pi = _pi_w(s, ia, torch)
    ref = orc.dyson(Pi.cpu().numpy(), s.V)
    W = ctx.dyson_w(Pi.clone(), V)
    err = float(np.max(np.abs(W.cpu().numpy() - ref)) / np.max(np.abs(ref)))
    assert err < 1e-10, err
    ctx.close()
