"""IR sparse-sampling tables for the imaginary-axis transforms.

Mirrors imag_axes_ft::ir::IR (src/numerics/imag_axes_ft/ir/ir_driver.hpp:23-76): lambda is
beta*wmax rounded up to {1e2, ..., 1e6} (:157-171) and the precision label selects the
eps of the table (:32-42).  Tables are the reference's HDF5 data converted to .npz
(coqui_amd/data/ir/convert_ir_tables.py); they are kept dimensionless here and rescaled
with beta inside libthcgw, as ir_driver.hpp:68-73 does.
"""
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "ir")
_PREC = {"high": "1e-15", "medium": "1e-10", "low": "1e-06"}
_LAMBDA = {100: "1e2", 1000: "1e3", 10000: "1e4", 100000: "1e5", 1000000: "1e6"}


def determine_lambda(lbda):
    """ir_driver.hpp:157-171."""
    for lam in (100, 1000, 10000, 100000, 1000000):
        if lbda <= lam:
            return lam
    return -1


def available_tables():
    return sorted(fn[:-4] for fn in os.listdir(_DATA) if fn.endswith(".npz"))


class IAFT:
    """Dimensionless IR tables plus beta.

    IAFT(beta, wmax, prec) follows the reference's table choice; IAFT.from_table(beta, name)
    picks a table by file name (e.g. "1e3.1e-15")."""

    def __init__(self, beta, wmax, prec="high"):
        if prec not in _PREC:
            raise ValueError("prec = %r is not acceptable; use high, medium or low" % prec)
        lam = determine_lambda(beta * wmax)
        if lam < 0:
            raise ValueError("invalid lambda for beta*wmax = %g" % (beta * wmax))
        self._load(beta, "%s.%s" % (_LAMBDA[lam], _PREC[prec]))
        self.wmax = lam / beta
        self.prec = prec

    @classmethod
    def from_table(cls, beta, name):
        obj = cls.__new__(cls)
        obj._load(beta, name)
        return obj

    @classmethod
    def with_ntau(cls, beta, nt):
        """The shipped table with exactly nt fermionic and bosonic tau points."""
        for name in available_tables():
            z = np.load(os.path.join(_DATA, name + ".npz"))
            if int(z["f_nt"]) == nt and int(z["b_nt"]) == nt:
                return cls.from_table(beta, name)
        raise ValueError("no IR table with nt = %d (available: %s)" % (nt, available_tables()))

    def _load(self, beta, name):
        path = os.path.join(_DATA, name + ".npz")
        if not os.path.exists(path):
            raise FileNotFoundError("IR table %s not found (available: %s)" % (name, available_tables()))
        z = np.load(path)
        self.name = name
        self.beta = float(beta)
        self.lambda_ = float(name.split(".")[0])
        self.nt_f = int(z["f_nt"])
        self.nw_f = int(z["f_nw"])
        self.nt_b = int(z["b_nt"])
        self.nw_b = int(z["b_nw"])
        if self.nt_f != self.nt_b:
            raise ValueError("nt_f != nt_b in table %s; the THC GW path assumes equal tau meshes" % name)
        self.tau_mesh = np.ascontiguousarray(z["f_tau_mesh"], np.float64)  # x in [-1, 1]
        self.wn_mesh_f = np.ascontiguousarray(z["f_wn_mesh"], np.int64)
        self.wn_mesh_b = np.ascontiguousarray(z["b_wn_mesh"], np.int64)
        c = lambda a: np.ascontiguousarray(a, np.complex128)
        self.Ttw_ff = c(z["f_Ttw"])
        self.Twt_ff = c(z["f_Twt"])
        self.Ttw_bb = c(z["b_Ttw"])
        self.Twt_bb = c(z["b_Twt"])
        self.T_beta_t_ff = c(z["f_T_betat"])

    @property
    def nt(self):
        return self.nt_f

    @property
    def tau(self):
        """tau = (x + 1) beta / 2 (qp_scf_common.cpp:96-98)."""
        return (self.tau_mesh + 1.0) * self.beta / 2.0

    @property
    def nt_half(self):
        return (self.nt + 1) // 2

    @property
    def nw_half(self):
        return (self.nw_b + 1) // 2
