"""The IR table generator (coqui_amd/data/ir/gen_ir_tables.py + sve_logistic.cpp).

* Regenerating the reference's shipped lambda = 1e2, eps = 1e-15 table end to end (SVE in
  binary128, sampling points, transforms) reproduces its meshes and transforms.  The
  shipped tables were written by sparse_ir 1.1.7 (gen_sparse_ir.py:9-105), so this pins the
  sampling rules (roots of u_L, sign changes of uhat, their rounding to Matsubara indices)
  and the normalisation of the transforms against independent data.
* The generated 80- and 60-point tables (1e3.5e-18, 1e3.1e-12) pass the checks that the
  reference applies to its own tables (test_iaft.cpp:40-43: Ttw Twt = I) and the single-pole
  closed form, in the reference's rescaled convention (ir_driver.hpp:68-73).
"""
import os
import tempfile

import numpy as np
import pytest
from scipy.special import expit

from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IRDIR = os.path.join(ROOT, "coqui_amd", "data", "ir")


def _gen():
    import importlib.util

    spec = importlib.util.spec_from_file_location("gen_ir_tables", os.path.join(IRDIR, "gen_ir_tables.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_regenerates_shipped_table():
    g = _gen()
    with tempfile.TemporaryDirectory() as d:
        sve = g.SVE.compute(100.0, exe=g.build_tool(d))
    # basis size of the shipped table: s_l >= 1e-15 s_0 for l < 40
    assert int(np.sum(sve.s >= 1e-15 * sve.s[0])) == 40
    dev = g.compare("1e2.1e-15", sve)
    for k in ("f_nt", "f_nw", "b_nt", "b_nw", "f_wn_mesh", "b_wn_mesh"):
        assert dev[k] is True, k
    assert dev["f_tau_mesh"] < 1e-14 and dev["b_tau_mesh"] < 1e-14
    for k in ("f_Ttw", "f_Twt", "b_Ttw", "b_Twt", "f_T_betat", "f_T_zerot", "f_Ttt_bf", "b_Ttt_fb", "f_Tct", "b_Tct"):
        assert dev[k] < 1e-11, (k, dev[k])


def test_singular_functions_orthonormal():
    """u_l orthonormal on [-1, 1] (Gauss-Legendre quadrature on the expansion's own segments)."""
    g = _gen()
    with tempfile.TemporaryDirectory() as d:
        sve = g.SVE.compute(100.0, exe=g.build_tool(d))
    t, w = np.polynomial.legendre.leggauss(48)
    xs, ws = [], []
    for a, b in zip(sve.knots[:-1], sve.knots[1:]):
        for sgn in (1.0, -1.0):
            xs.append(sgn * ((a + b) / 2 + (b - a) / 2 * t))
            ws.append((b - a) / 2 * w)
    x, wq = np.concatenate(xs), np.concatenate(ws)
    U = np.stack([sve.u(l, x) for l in range(45)])
    gram = (U * wq) @ U.T
    assert np.max(np.abs(gram - np.eye(45))) < 1e-12
    # parity alternates with decreasing singular value, starting even
    assert np.all(sve.parity[:45] == np.where(np.arange(45) % 2 == 0, 1, -1))


@pytest.mark.parametrize("table,nt,nwf,nwb", [("1e3.5e-18", 80, 80, 81), ("1e3.1e-12", 60, 60, 61)])
def test_generated_table_shapes(table, nt, nwf, nwb):
    z = np.load(os.path.join(IRDIR, table + ".npz"))
    assert int(z["f_nt"]) == nt and int(z["b_nt"]) == nt
    assert int(z["f_nw"]) == nwf and int(z["b_nw"]) == nwb
    x = z["f_tau_mesh"]
    assert np.all(np.diff(x) > 0) and -1 < x[0] and x[-1] < 1
    assert np.max(np.abs(x + x[::-1])) < 1e-15  # symmetric about tau = beta/2
    assert np.array_equal(z["f_wn_mesh"], -z["f_wn_mesh"][::-1]) and np.all(z["f_wn_mesh"] % 2 == 1)
    assert np.array_equal(z["b_wn_mesh"], -z["b_wn_mesh"][::-1]) and np.all(z["b_wn_mesh"] % 2 == 0)


@pytest.mark.parametrize("table", ["1e3.5e-18", "1e3.1e-12"])
def test_generated_identity(table):
    """test_iaft.cpp:40-43 on the generated tables: Ttw Twt = I."""
    tab = oracle.load_ir(1000.0, None, table=table)
    assert np.max(np.abs(tab["Ttw_ff"] @ tab["Twt_ff"] - np.eye(tab["nt"]))) < 1e-10
    assert np.max(np.abs(tab["Ttw_bb"] @ tab["Twt_bb"] - np.eye(tab["nt"]))) < 1e-9


@pytest.mark.parametrize("table,tol", [("1e3.5e-18", 1e-10), ("1e3.1e-12", 1e-8)])
def test_generated_single_pole(table, tol):
    """g(tau) = -exp(-e tau)/(1 + exp(-e beta)) maps to 1/(iw_n - e) for |e| <= wmax, and to
    g(beta^-) = -f(e) through T_beta; the tolerance follows the table's eps."""
    beta = 1000.0
    tab = oracle.load_ir(beta, None, table=table)
    t = tab["tau"]
    wmax = float(table.split(".")[0]) / beta
    for e in (-0.9 * wmax, -0.3, -0.004, 0.02, 0.7, 0.95 * wmax):
        gt = -np.exp(-e * t) / (1 + np.exp(-e * beta)) if e > 0 else -np.exp(e * (beta - t)) / (1 + np.exp(e * beta))
        gw = oracle.transform(tab["Twt_ff"], gt[:, None].astype(complex))[:, 0]
        iw = 1j * tab["wn_f"] * np.pi / beta
        assert np.max(np.abs(gw - 1.0 / (iw - e))) < tol * beta, (table, e)
        back = oracle.transform(tab["Ttw_ff"], (1.0 / (iw - e))[:, None])[:, 0]
        assert np.max(np.abs(back - gt)) < tol
        gb = oracle.tau_to_beta(tab, gt[:, None].astype(complex))[0]
        assert abs(gb + expit(-beta * e)) < tol  # f(e) = 1/(1 + exp(beta e))
