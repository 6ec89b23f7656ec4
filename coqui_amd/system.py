"""Seeded synthetic THC systems.

One generator serves the GPU path, the CPU oracle and the benchmark, so both sides of a
parity check see the same bytes.  A system is:

  * a Gamma-centred kgrid[0] x kgrid[1] x kgrid[2] k-mesh (lattice vectors = identity),
    ns = 1, no symmetry, div_treatment = ignore_g0;
  * H(k) = U diag(e) U^H with U Haar-random (QR of a complex Gaussian matrix with the
    phase of R's diagonal moved into Q) and e uniform in [-1, 1] Ha, sorted; mu = 0.
    The MO coefficients are C = U and the energies E = e;
  * X^k[mu, i] complex Gaussian with each row normalised over i;
  * V_q = Z_q Z_q^H / N_mu with Z_q complex Gaussian N_mu x N_mu.  V_q is used as the
    reference's thc_reader_t::Z(iq) (thc_reader_t.hpp:228-230), i.e. the N_k factor is
    taken to be included in it.

Seeds are 20261202 + 1000 * system_id.

With trev=True the system is made time-reversal symmetric after it is drawn: for every
k-point whose partner -k (modulo the grid) is a different point, the partner's C, E, X and
Z are the complex conjugates of the point's own; at the points with -k = k (time-reversal
invariant momenta) C is replaced by a real orthogonal matrix (QR of the real part) and X
and Z by their real parts.  Then H(-k) = conj H(k), X(-k) = conj X(k), V(-q) = conj V(q),
and so Pi(-q) = conj Pi(q), W(-q) = conj W(q) and Sigma(-k) = conj Sigma(k).

With c4=True (kgrid[0] == kgrid[1]) the system has a fourfold rotation S(i, j, l) =
(-j, i, l) acting on the THC points as a fixed permutation pi of order 4: on every orbit
of S the first point is drawn freely (X and Z averaged over the orbit's stabiliser) and
the point S^a k gets H(k), X(S^a k) = X(k)[pi_a] and Z(S^a k) = Z(k)[pi_a][:, pi_a], so
V, Pi and W obey A(S^a q) = A(q)[pi_a][:, pi_a].  System.qsym is then (src, perm): the
orbit representative of every q and the index map pi_a (nk x nmu).
"""
from dataclasses import dataclass, field

import numpy as np

BASE_SEED = 20261202


@dataclass
class System:
    kgrid: tuple
    no: int
    nmu: int
    seed: int
    C: np.ndarray = field(repr=False)    # (nk, no, no) MO coefficients, column a = orbital
    E: np.ndarray = field(repr=False)    # (nk, no) MO energies
    X: np.ndarray = field(repr=False)    # (nk, nmu, no)
    V: np.ndarray = field(repr=False)    # (nk, nmu, nmu), or None when built on the device
    Z: np.ndarray = field(repr=False)    # (nk, nmu, nmu) factor of V
    mu: float = 0.0
    qsym: tuple = field(default=None, repr=False)  # (src (nk), perm (nk, nmu)) with c4=True

    @property
    def nk(self):
        return int(np.prod(self.kgrid))

    @property
    def H(self):
        """H(k) = C diag(E) C^H: the primary-basis Hamiltonian used as Vhf in the QP equation."""
        return np.einsum("kia,ka,kja->kij", self.C, self.E, self.C.conj())


def _cgauss(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)


def minus_k(kgrid):
    """Index of -k (modulo the grid) for every k of the C-order kgrid."""
    n1, n2, n3 = kgrid
    out = np.empty(n1 * n2 * n3, np.int64)
    for i in range(n1):
        for j in range(n2):
            for l in range(n3):
                out[(i * n2 + j) * n3 + l] = (((-i) % n1) * n2 + (-j) % n2) * n3 + (-l) % n3
    return out


def c4_orbits(kgrid):
    """For every k of the grid: (representative r, power a) with k = S^a r, S(i, j, l) = (-j, i, l)."""
    n1, n2, n3 = kgrid
    assert n1 == n2, "the fourfold rotation needs kgrid[0] == kgrid[1]"
    idx = lambda i, j, l: (i * n2 + j) * n3 + l
    nk = n1 * n2 * n3
    rep = [-1] * nk
    power = [0] * nk
    for k in range(nk):
        if rep[k] >= 0:
            continue
        i, j, l = k // (n2 * n3), (k // n3) % n2, k % n3
        for a in range(4):
            t = idx(i, j, l)
            if rep[t] < 0:
                rep[t], power[t] = k, a
            i, j = (-j) % n1, i
    return np.array(rep), np.array(power)


def _perm_c4(nmu, rng):
    """A permutation of order 4 (4-cycles, 2-cycles and fixed points) as index maps pi^a, a = 0..3."""
    order = rng.permutation(nmu)
    base = np.arange(nmu)
    p = base.copy()
    pos = 0
    sizes = [4, 2, 1]
    while pos < nmu:
        c = sizes[rng.integers(0, 3)] if nmu - pos >= 4 else 1
        cyc = order[pos:pos + c]
        p[cyc] = np.roll(cyc, 1)
        pos += c
    maps = [base]
    for _ in range(3):
        maps.append(p[maps[-1]])
    return maps


def make_system(kgrid, no, nmu, system_id=0, with_v=True, trev=False, c4=False):
    """Build a synthetic system.  with_v=False skips V (the caller forms Z Z^H / N_mu itself);
    trev=True makes it time-reversal symmetric and c4=True symmetric under the fourfold
    rotation (see the module docstring)."""
    kgrid = tuple(int(v) for v in kgrid)
    nk = int(np.prod(kgrid))
    seed = BASE_SEED + 1000 * int(system_id)
    rng = np.random.default_rng(seed)
    A = _cgauss(rng, (nk, no, no))
    Q, R = np.linalg.qr(A)
    d = np.diagonal(R, axis1=1, axis2=2)
    U = Q * (d / np.abs(d))[:, None, :]
    e = rng.uniform(-1.0, 1.0, (nk, no))
    order = np.argsort(e, axis=1)
    E = np.take_along_axis(e, order, axis=1)
    C = np.ascontiguousarray(np.take_along_axis(U, order[:, None, :], axis=2))
    X = _cgauss(rng, (nk, nmu, no))
    X /= np.linalg.norm(X, axis=2, keepdims=True)
    Z = _cgauss(rng, (nk, nmu, nmu))
    if trev:
        mk = minus_k(kgrid)
        for k in range(nk):
            p = int(mk[k])
            if p == k:
                q, r = np.linalg.qr(C[k].real)
                C[k] = q * np.sign(np.diagonal(r))[None, :]
                X[k] = X[k].real / np.linalg.norm(X[k].real, axis=1, keepdims=True)
                Z[k] = Z[k].real
            elif p > k:
                C[p], E[p], X[p], Z[p] = C[k].conj(), E[k], X[k].conj(), Z[k].conj()
    qsym = None
    if c4:
        rep, power = c4_orbits(kgrid)
        maps = _perm_c4(nmu, rng)
        V0 = np.matmul(Z, np.conj(np.swapaxes(Z, 1, 2))) / nmu
        Vs = np.empty_like(V0)
        for r in sorted(set(rep.tolist())):
            members = [k for k in range(nk) if rep[k] == r]
            s_ = len(members)  # the stabiliser of r is generated by S^s_
            stab = [maps[(s_ * b) % 4] for b in range(4 // s_)]
            X[r] = sum(X[r][m] for m in stab) / len(stab)
            X[r] /= np.linalg.norm(X[r], axis=1, keepdims=True)
            Vs[r] = sum(V0[r][np.ix_(m, m)] for m in stab) / len(stab)
            for k in members:
                if k != r:
                    m = maps[power[k]]
                    C[k], E[k], X[k] = C[r], E[r], X[r][m]
                    Vs[k] = Vs[r][np.ix_(m, m)]
        perm = np.stack([maps[power[k]] for k in range(nk)]).astype(np.int32)
        qsym = (rep.astype(np.int32), perm)
        V = Vs if with_v else None
        Z = None
    else:
        V = None
        if with_v:
            V = np.matmul(Z, np.conj(np.swapaxes(Z, 1, 2))) / nmu
    return System(kgrid=kgrid, no=no, nmu=nmu, seed=seed, C=C, E=np.ascontiguousarray(E), X=np.ascontiguousarray(X),
                  V=V, Z=Z, qsym=qsym)


def batch_specs(n=24):
    """The batched workload: n systems with N_orb ~ U{16..96}, N_mu = 8 N_orb and a k-grid of
    2^3, 3^3 or 4^3 points, drawn from the base seed; system i uses seed id i.  Returns
    [(kgrid, N_orb, N_mu, system_id)]."""
    rng = np.random.default_rng(BASE_SEED)
    out = []
    for i in range(n):
        no = int(rng.integers(16, 97))
        n1 = int(rng.choice([2, 3, 4]))
        out.append(((n1, n1, n1), no, 8 * no, i))
    return out
