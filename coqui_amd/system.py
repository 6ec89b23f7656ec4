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
    c4_maps: list = field(default=None, repr=False)  # pi^a, a = 0..3 (the THC-point maps of S^a) with c4=True

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
                  V=V, Z=Z, qsym=qsym, c4_maps=maps if c4 else None)


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


def _c4_apply(kgrid, k, b):
    """Grid index of S^b k, S(i, j, l) = (-j, i, l)."""
    n1, n2, n3 = kgrid
    i, j, l = k // (n2 * n3), (k // n3) % n2, k % n3
    for _ in range(b % 4):
        i, j = (-j) % n1, i
    return (i * n2 + j) * n3 + l


def make_system_bz(kgrid, no, nmu, system_id=0, kind="c4"):
    """A symmetric system and its Brillouin-zone symmetry tables in the reference's layout
    (bz_symmetry.hpp:497-576): k-points listed with the irreducible ones first and the
    time-reversal partners last.

    kind="trev": the time-reversal symmetric system of make_system(trev=True).  The
    irreducible k are those with k <= -k (C order); the others are listed after them as
    time-reversal partners (kp_trev = 1).  One operation (the identity); q-points map to
    the irreducible ones with qp_trev where -q is irreducible.

    kind="c4": the fourfold-rotation system of make_system(c4=True), with orbital
    representations on the little groups: at an orbit representative r whose stabiliser is
    generated by T = S^s (order m = 4/s), orbital i carries the phase w_i = exp(2 pi i c_i / m)
    under T, so that X_r[pi_T] = X_r diag(w), H(r) commutes with diag(w), and the orbital
    rotation D of an operation that maps r into a point of its own orbit stored through
    another power of S is diag(w)^b, not the identity.  The operations are S^b (b = 0..3,
    those some q uses); q maps into the irreducible set through S^b when S^b q is its
    representative.

    Returns (system, bz): system is in C order (for full-grid checks); bz holds
      order (nk): grid index of each listed k-point (the q-points are listed the same way),
      nk_ibz, nq_ibz, nk_trev_pairs, kp_to_ibz, kp_trev, kp_trev_pair (nk),
      qk_to_k2 (nk, nk): [q, a] = b with k_b = k_a - q_q, qminus (nk), qp_to_ibz, qp_trev (nk),
      qsymms (nsym), nq_per_s (nsym), Qs (nsym, nk; -1 past nq_per_s), ks_to_k (nsym, nk_ibz),
      D (nsym, nk_ibz, no, no): the orbital rotation of the reference's symmetry_rotation,
      Sigma_ij(k) += sum_ab conj(D_ai) Sigma_ab(k S) D_bj."""
    kgrid = tuple(int(v) for v in kgrid)
    nk = int(np.prod(kgrid))
    mk = minus_k(kgrid)
    if kind == "trev":
        s = make_system(kgrid, no, nmu, system_id=system_id, trev=True)
        ibz = [k for k in range(nk) if k <= mk[k]]
        rest = [k for k in range(nk) if k > mk[k]]
        order = np.array(ibz + rest, np.int64)
        pos = np.empty(nk, np.int64)
        pos[order] = np.arange(nk)
        nk_ibz = len(ibz)
        kp_to_ibz = np.array([pos[min(g, mk[g])] for g in order], np.int32)
        kp_trev = np.array([int(g > mk[g]) for g in order], np.int32)
        kp_trev_pair = np.array([pos[mk[g]] if mk[g] != g else -1 for g in order], np.int32)
        qsymms = [0]
        qp_symm = np.zeros(nk, np.int64)
        ks_to_k = np.arange(nk_ibz, dtype=np.int32)[None, :]
        D = np.broadcast_to(np.eye(no, dtype=np.complex128), (1, nk_ibz, no, no)).copy()
        qp_to_ibz, qp_trev = kp_to_ibz.copy(), kp_trev.copy()
    elif kind == "c4":
        s = make_system(kgrid, no, nmu, system_id=system_id, c4=True)
        rep, power = c4_orbits(kgrid)
        maps = s.c4_maps
        rng = np.random.default_rng(s.seed + 7)
        H = s.H
        C, E, X = s.C.copy(), s.E.copy(), s.X.copy()
        reps = sorted(set(rep.tolist()))
        for r in reps:
            members = [k for k in range(nk) if rep[k] == r]
            sz = len(members)
            if sz == 4:
                continue
            m = 4 // sz
            c = rng.integers(0, m, no)
            c[0] = 0  # a point fixed by T keeps a non-zero row
            w = np.exp(2j * np.pi * c / m)
            X0 = _cgauss(rng, (nmu, no))
            Xr = sum(X0[maps[(sz * b) % 4]] * (w ** (-b))[None, :] for b in range(m)) / m
            Xr /= np.linalg.norm(Xr, axis=1, keepdims=True)
            Hr = H[r] * (c[:, None] == c[None, :])
            e, u = np.linalg.eigh(Hr)
            for k in members:
                C[k], E[k], X[k] = u, e, Xr[maps[power[k]]]
        s.C, s.E, s.X = C, np.ascontiguousarray(E), np.ascontiguousarray(X)
        rest = [k for k in range(nk) if rep[k] != k]
        order = np.array(reps + rest, np.int64)
        pos = np.empty(nk, np.int64)
        pos[order] = np.arange(nk)
        nk_ibz = len(reps)
        kp_to_ibz = np.array([pos[rep[g]] for g in order], np.int32)
        kp_trev = np.zeros(nk, np.int32)
        kp_trev_pair = np.full(nk, -1, np.int32)
        qp_symm_all = np.array([(4 - power[g]) % 4 for g in order], np.int64)
        qsymms = sorted(set(qp_symm_all.tolist()))
        qp_symm = np.array([qsymms.index(b) for b in qp_symm_all], np.int64)
        ks_to_k = np.array([[pos[_c4_apply(kgrid, order[ik], b)] for ik in range(nk_ibz)] for b in qsymms], np.int32)
        D = np.empty((len(qsymms), nk_ibz, no, no), np.complex128)
        for isym, b in enumerate(qsymms):
            for ik in range(nk_ibz):
                r = order[ik]
                Xks = X[order[ks_to_k[isym, ik]]]
                target = X[r][maps[b]]
                d, *_ = np.linalg.lstsq(Xks, target, rcond=None)
                assert np.abs(Xks @ d - target).max() < 1e-10, "orbital rotation does not close"
                D[isym, ik] = d
        qp_to_ibz, qp_trev = kp_to_ibz.copy(), kp_trev.copy()
    else:
        raise ValueError("kind is 'c4' or 'trev'")
    nsym = len(qsymms)
    Qs = np.full((nsym, nk), -1, np.int32)
    nq_per_s = np.zeros(nsym, np.int32)
    for q in range(nk):
        i = qp_symm[q]
        Qs[i, nq_per_s[i]] = q
        nq_per_s[i] += 1
    n1, n2, n3 = kgrid
    qk_to_k2 = np.empty((nk, nk), np.int32)
    for q in range(nk):
        gq = order[q]
        qi, qj, ql = gq // (n2 * n3), (gq // n3) % n2, gq % n3
        for a in range(nk):
            ga = order[a]
            ai, aj, al = ga // (n2 * n3), (ga // n3) % n2, ga % n3
            qk_to_k2[q, a] = pos[(((ai - qi) % n1) * n2 + (aj - qj) % n2) * n3 + (al - ql) % n3]
    bz = dict(order=order.astype(np.int32), nk_ibz=nk_ibz, nq_ibz=nk_ibz, nk_trev_pairs=int(kp_trev.sum()),
              kp_to_ibz=kp_to_ibz, kp_trev=kp_trev, kp_trev_pair=kp_trev_pair, qk_to_k2=qk_to_k2,
              qminus=np.array([pos[mk[g]] for g in order], np.int32), qp_to_ibz=qp_to_ibz, qp_trev=qp_trev,
              qsymms=np.array(qsymms, np.int32), nq_per_s=nq_per_s, Qs=Qs, ks_to_k=np.ascontiguousarray(ks_to_k),
              D=D)
    return s, bz
