"""Intermediate-representation (IR) tables built from the singular value expansion of the
logistic kernel, in the layout of the reference's tables.

The reference precomputes its tables offline with the sparse_ir library
(src/numerics/imag_axes_ft/ir/data/gen_sparse_ir.py:9-105) and ships them for
lambda in {1e2, ..., 1e6} and eps in {1e-15, 1e-10, 1e-6} (nt = 40/72/104/... at 1e-15).
This module builds tables of any (lambda, eps) from the expansion computed by
sve_logistic.cpp (binary128 QR + Jacobi SVD of the discretised kernel), following the
published construction (Shinaoka et al., PRB 96, 035147 (2017); Chikano et al., CPC 240,
181 (2019)) and the sampling rules of sparse_ir 1.1.7, the version that wrote the shipped
tables:

* basis size L = #{l : s_l >= eps s_0};
* tau sampling points: the L roots of u_L, the first singular function outside the basis;
* Matsubara sampling points: the sign changes of uhat_l' over the integers, with l' = L
  rounded up to even (fermions) or odd (bosons), each taken at the lower of the two
  neighbouring Matsubara indices, mirrored to negative frequencies; bosons add n = 0;
* transforms (gen_sparse_ir.py:34-65): Ttl = u_l(x_t), Twl = uhat_l(n) with
  uhat_l(n) = 2^-1/2 int_{-1}^{1} exp(i pi n (x + 1) / 2) u_l(x) dx, Ttw = Ttl pinv(Twl),
  Twt = Twl pinv(Ttl), Ttt = Ttl pinv(Ttl), T_beta = u(1) pinv(Ttl), T_zero = u(-1) pinv(Ttl),
  Tct = pinv(Ttl); all dimensionless (coqui_amd.iaft rescales with beta as
  ir_driver.hpp:68-73).

Regenerating a shipped table reproduces its meshes (Matsubara indices exactly, tau points to
~1e-16) and transforms (to ~1e-12 relative); tests/test_ir_generator.py checks this.

    python coqui_amd/data/ir/gen_ir_tables.py 1e3 5e-18      # writes 1e3.5e-18.npz (80 points)
    python coqui_amd/data/ir/gen_ir_tables.py 1e3 1e-12      # writes 1e3.1e-12.npz (60 points)
    python coqui_amd/data/ir/gen_ir_tables.py --compare 1e3.1e-15   # against the shipped table
"""
import argparse
import os
import subprocess
import sys
import tempfile

import numpy as np
from numpy.polynomial import legendre as npl
from scipy.optimize import brentq
from scipy.special import spherical_jn

HERE = os.path.dirname(os.path.abspath(__file__))
VERSION = "coqui_amd-sve-1"


def build_tool(out_dir=None):
    """Compile sve_logistic.cpp (g++, libquadmath) and return the executable's path."""
    out_dir = out_dir or os.path.join(HERE, "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "sve_logistic")
    src = os.path.join(HERE, "sve_logistic.cpp")
    if not os.path.exists(exe) or os.path.getmtime(exe) < os.path.getmtime(src):
        subprocess.run(["g++", "-O2", "-fopenmp", src, "-lquadmath", "-o", exe], check=True)
    return exe


class SVE:
    """Singular functions u_l(x) of the logistic kernel on [-1, 1], piecewise Legendre."""

    def __init__(self, path):
        b = open(path, "rb").read()
        nseg, ng, nf = (int(v) for v in np.frombuffer(b[:12], np.int32))
        off = 12
        self.knots = np.frombuffer(b[off:off + 8 * (nseg + 1)], np.float64).copy()
        off += 8 * (nseg + 1)
        rec = 2 + nseg * ng
        d = np.frombuffer(b[off:off + 8 * rec * nf], np.float64).reshape(nf, rec)
        self.parity = d[:, 0].astype(int)
        self.s = d[:, 1].copy()
        self.coef = d[:, 2:].reshape(nf, nseg, ng).copy()
        # sign convention u_l(1) > 0 (the value at x = 1 is the sum of the last segment's
        # Legendre coefficients); it only shows in Tct = pinv(Ttl), whose rows follow u_l
        flip = self.coef[:, -1, :].sum(axis=1) < 0
        self.coef[flip] *= -1.0

    @classmethod
    def compute(cls, lam, n_gauss=32, exe=None):
        exe = exe or build_tool()
        with tempfile.TemporaryDirectory() as d:
            out = os.path.join(d, "sve.bin")
            subprocess.run([exe, repr(float(lam)), out, str(n_gauss)], check=True, stderr=subprocess.DEVNULL)
            return cls(out)

    def u(self, l, x):
        """u_l(x), orthonormal on [-1, 1]."""
        x = np.atleast_1d(np.asarray(x, np.float64))
        ax = np.abs(x)
        seg = np.clip(np.searchsorted(self.knots, ax, side="right") - 1, 0, len(self.knots) - 2)
        a, b = self.knots[seg], self.knots[seg + 1]
        t = (2.0 * ax - a - b) / (b - a)
        out = np.empty_like(ax)
        for sg in np.unique(seg):
            m = seg == sg
            out[m] = npl.legval(t[m], self.coef[l, sg])
        if self.parity[l] < 0:
            out *= np.sign(x)
        return out / np.sqrt(2.0)

    def uhat(self, ls, n):
        """uhat[i, j] = 2^-1/2 int_{-1}^{1} exp(i pi n_j (x + 1)/2) u_{ls[i]}(x) dx.

        Per segment [a, b] of [0, 1] with x = m + h t: int P_k(t) exp(i w h t) dt = 2 i^k j_k(w h);
        the even (odd) extension to [-1, 0] doubles the real (imaginary) part."""
        n = np.asarray(n, np.float64)
        w = np.pi * n / 2.0
        ls = list(ls)
        acc = np.zeros((len(ls), n.size), np.complex128)
        ng = self.coef.shape[2]
        ik = np.array([2.0 * (1j ** k) for k in range(ng)])
        for sg in range(len(self.knots) - 1):
            a, b = self.knots[sg], self.knots[sg + 1]
            m, h = (a + b) / 2.0, (b - a) / 2.0
            J = np.stack([spherical_jn(k, w * h) for k in range(ng)]) * ik[:, None]
            acc += (self.coef[ls, sg, :] @ J) * (h * np.exp(1j * w * m))[None, :]
        par = self.parity[ls][:, None]
        out = np.where(par > 0, 2.0 * acc.real + 0j, 2j * acc.imag) / np.sqrt(2.0)
        return out * np.exp(1j * np.pi * n / 2.0)[None, :] / np.sqrt(2.0)

    def roots(self, l):
        """All roots of u_l in (-1, 1), ascending."""
        f = lambda z: float(self.u(l, z)[0])
        xs = []
        for sg in range(len(self.knots) - 1):
            g = np.linspace(self.knots[sg], self.knots[sg + 1], 257)
            v = self.u(l, g)
            for i in range(len(g) - 1):
                if v[i] == 0.0 and g[i] > 0:
                    xs.append(g[i])
                elif v[i] * v[i + 1] < 0:
                    xs.append(brentq(f, g[i], g[i + 1], xtol=1e-17, rtol=1e-15, maxiter=200))
        xs = np.unique(np.array([x for x in xs if x > 0]))
        mid = [0.0] if self.parity[l] < 0 else []
        return np.concatenate([-xs[::-1], mid, xs])


def _sign_changes(sve, l, zeta, mmax):
    """Positive Matsubara indices n = 2m + zeta at which the non-vanishing part of uhat_l
    changes sign between m and m + 1 (the lower index is returned)."""
    # all integers up to 4096, then a geometric grid (ratio 1.002) with integer bisection
    dense = np.arange(0, min(4096, mmax))
    geo = np.unique(np.round(4096 * 1.002 ** np.arange(0, int(np.log(max(mmax, 4097) / 4096) / np.log(1.002)) + 2)))
    grid = np.unique(np.concatenate([dense, geo.astype(np.int64)]))
    part = lambda m: _part(sve, l, zeta, m)
    v = part(grid)
    out = []
    for i in np.nonzero(np.signbit(v[:-1]) != np.signbit(v[1:]))[0]:
        a, b = int(grid[i]), int(grid[i + 1])
        fa = v[i]
        while b - a > 1:
            c = (a + b) // 2
            fc = part(np.array([c]))[0]
            if np.signbit(fc) == np.signbit(fa):
                a, fa = c, fc
            else:
                b = c
        out.append(2 * a + zeta)
    return np.array(out, np.int64)


def _part(sve, l, zeta, m):
    m = np.asarray(m, np.int64)
    vals = []
    for i in range(0, m.size, 8192):
        vals.append(sve.uhat([l], 2 * m[i:i + 8192] + zeta)[0])
    f = np.concatenate(vals)
    # even u: uhat is imaginary at odd n, real at even n; odd u the other way round
    return f.imag if ((sve.parity[l] > 0) == (zeta == 1)) else f.real


def make_table(sve, lam, eps):
    s = sve.s
    L = int(np.sum(s >= eps * s[0]))
    if L + 2 > s.size:
        raise ValueError("expansion too short for eps = %g" % eps)
    x = sve.roots(L)
    if x.size != L:
        raise RuntimeError("u_%d has %d roots, expected %d" % (L, x.size, L))
    ls = np.arange(L)
    Ttl = np.stack([sve.u(l, x) for l in ls], axis=1)  # (nt, L)
    pinvT = np.linalg.pinv(Ttl)
    mmax = int(40 * lam)
    out = {"version": np.array(VERSION)}
    for stat, zeta in (("f", 1), ("b", 0)):
        lreq = L + (1 if (stat == "f" and L % 2 == 1) or (stat == "b" and L % 2 == 0) else 0)
        pos = _sign_changes(sve, lreq, zeta, mmax)
        if stat == "b":
            pos = pos[pos > 0]
            wn = np.concatenate([-pos[::-1], [0], pos])
        else:
            wn = np.concatenate([-pos[::-1], pos])
        if wn.size != lreq:
            raise RuntimeError("%s: %d Matsubara points, expected %d" % (stat, wn.size, lreq))
        Twl = sve.uhat(ls, wn).T  # (nw, L)
        out[stat + "_nt"] = np.int64(L)
        out[stat + "_nw"] = np.int64(wn.size)
        out[stat + "_tau_mesh"] = x.copy()
        out[stat + "_wn_mesh"] = wn.astype(np.int64)
        out[stat + "_Ttw"] = Ttl @ np.linalg.pinv(Twl)
        out[stat + "_Twt"] = Twl @ pinvT
        out[stat + "_Tct"] = pinvT.astype(np.complex128)
        if stat == "f":
            out["f_Ttt_bf"] = (Ttl @ pinvT).astype(np.complex128)
            out["f_T_betat"] = (np.array([sve.u(l, 1.0)[0] for l in ls]) @ pinvT).astype(np.complex128)
            out["f_T_zerot"] = (np.array([sve.u(l, -1.0)[0] for l in ls]) @ pinvT).astype(np.complex128)
        else:
            out["b_Ttt_fb"] = (Ttl @ pinvT).astype(np.complex128)
    return out


def compare(table, sve=None):
    """Max deviations of a regenerated table from the shipped one of the same name."""
    lam, eps = (float(v) for v in table.split(".", 1))
    ref = np.load(os.path.join(HERE, table + ".npz"))
    got = make_table(sve or SVE.compute(lam), lam, eps)
    rel = lambda a, b: float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    res = {}
    for k in ref.files:
        if k == "version":
            continue
        a, b = got[k], ref[k]
        if np.ndim(b) == 0 or b.dtype.kind == "i":
            res[k] = bool(np.array_equal(a, b))
        elif k.endswith("tau_mesh"):
            res[k] = float(np.max(np.abs(a - b)))
        else:
            res[k] = rel(a, b)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("lam", nargs="?")
    ap.add_argument("eps", nargs="?")
    ap.add_argument("--compare", default=None, help="regenerate a shipped table and print the deviations")
    ap.add_argument("--out", default=HERE)
    args = ap.parse_args()
    if args.compare:
        for k, v in compare(args.compare).items():
            print("%-12s %s" % (k, v))
        return
    lam, eps = float(args.lam), float(args.eps)
    t = make_table(SVE.compute(lam), lam, eps)
    name = "%s.%s" % (args.lam, args.eps)
    np.savez_compressed(os.path.join(args.out, name + ".npz"), **t)
    print(name, int(t["f_nt"]), int(t["f_nw"]), int(t["b_nt"]), int(t["b_nw"]))


if __name__ == "__main__":
    sys.exit(main())
