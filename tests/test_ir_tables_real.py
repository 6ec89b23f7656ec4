"""The particle-hole folded tau <-> iW tables are real up to rounding.

The transform kernel applies only Re T when max |Im T| <= 1e-12 max |Re T| (checked when a
context is created; the complex product is used otherwise).  Folding adds the columns of
tau and beta - tau (IAFT.icc:62-70) and the rows of iW_n and -iW_n (IAFT.icc:117-125): for
bosonic frequencies e^{i W tau} + e^{i W (beta - tau)} = 2 cos(W tau), so the imaginary
parts cancel, and a table that breaks this fails here.  The fold is restated from the
reference, independent of the library's own fold on the host.
"""
import numpy as np
import pytest

from oracle import oracle


def fold(tab):
    T, Ti = tab["Twt_bb"], tab["Ttw_bb"]
    nwb, nt = T.shape
    nth = nt // 2 if nt % 2 == 0 else nt // 2 + 1
    nwh = nwb // 2 if nwb % 2 == 0 else nwb // 2 + 1
    Tp = np.zeros((nwh, nth), complex)
    Tm = np.zeros((nth, nwh), complex)
    for n in range(nwh):
        iw = nwb // 2 + n
        for it in range(nth):
            imt = nt - it - 1
            Tp[n, it] = T[iw, it] if it == imt else T[iw, it] + T[iw, imt]
    for it in range(nth):
        for n in range(nwh):
            iw, imw = nwb // 2 + n, nwb // 2 - n
            Tm[it, n] = Ti[it, iw] if iw == imw else Ti[it, iw] + Ti[it, imw]
    return Tp, Tm


@pytest.mark.parametrize("table", ["1e2.1e-15", "1e3.1e-15", "1e4.1e-15", "1e3.5e-18", "1e3.1e-12"])
def test_folded_tables_real(table):
    Tp, Tm = fold(oracle.load_ir(1000.0, None, table=table))
    for T in (Tp, Tm):
        ratio = np.abs(T.imag).max() / np.abs(T.real).max()
        assert ratio <= 1e-12, (table, ratio)


def test_unfolded_tables_are_complex():
    """The unfolded tables are genuinely complex: the check above is not vacuous."""
    tab = oracle.load_ir(1000.0, None, table="1e3.1e-15")
    T = tab["Twt_bb"]
    assert np.abs(T.imag).max() > 1e-3 * np.abs(T.real).max()
