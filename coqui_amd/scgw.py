"""Self-consistent GW loop on one GPU (the order of scf_loop, src/methods/SCF/scf_driver.cpp:96-170).

Per iteration:  [THC-HF: F = H0 + J - K of the density matrix (optional)]  ->  update_w +
evaluate (chi0 -> tau->iW -> Dyson W -> iW->tau -> Sigma)  ->  hermitize  ->  damping of
Sigma (the reference's solve_iterative with a fixed mixing)  ->  Dyson for G with the
chemical potential search (update_G)  ->  Galitskii-Migdal correlation energy.

The Dyson step is pluggable:
  dyson="direct"   the library's own iteration (thcgw_gw_iteration, with the context's Dyson
                   form and precision flags);
  dyson="learned"  the stages one by one with the Dyson step of a LearnedDyson object:
                   direct Hermitian solves that record training pairs in the first
                   `train_iters` iterations, then training, then Krylov solves with the
                   learned low-rank preconditioner (each accepted only at a true FP64
                   residual <= tol, else solved directly).
"""
import time

import numpy as np

from .thcgw import LearnedDyson


class ScGW:
    def __init__(self, ctx, system, nel_target=None, mixing=1.0, with_hf=False):
        self.ctx = ctx
        self.s = system
        self.torch = ctx._torch
        self.mixing = float(mixing)
        self.with_hf = bool(with_hf)
        t = self.torch
        self.H = t.from_numpy(np.ascontiguousarray(system.H[None], np.complex128)).cuda()  # (1, nk, no, no)
        self.X = t.from_numpy(system.X).cuda()
        self.V = t.from_numpy(system.V).cuda()
        self.mu = 0.0
        self.Sigma = t.zeros((ctx.nt, 1, ctx.nk, ctx.no, ctx.no), dtype=t.complex128, device="cuda")
        self.F = self.H.clone()
        if nel_target is None:
            nel_target = ctx.nelec(self.F, self.Sigma, 0.0)
        self.nel_target = float(nel_target)
        self.G, self.Dm = ctx.dyson_g(self.F, self.Sigma, self.mu)
        self.history = []
        self._big = None

    def _buffers(self):
        """Pi(tau) / W(tau) (same size, reused) and Pi(iW) / W(iW) for the stage-wise path."""
        if self._big is None:
            t, c = self.torch, self.ctx
            a = t.empty((c.nt_half, c.nk, c.nmu, c.nmu), dtype=t.complex128, device="cuda")
            w = t.empty((c.nw_half, c.nk, c.nmu, c.nmu), dtype=t.complex128, device="cuda")
            self._big = (a, w)
        return self._big

    def _sigma_direct(self):
        S, W, _ = self.ctx.gw_iteration(self.G, self.X, self.V, with_ec=False, hermitize=True)
        del W
        return S, None

    def _sigma_learned(self, lri, krylov, record):
        c = self.ctx
        a, w = self._buffers()
        c.pi(self.G, self.X, out=a)
        c.tau_to_w(a, out=w)
        st = lri.dyson(w, self.V, krylov=krylov, record=record)
        Wqt = a.view(c.nk, c.nt_half, c.nmu, c.nmu)  # Pi(tau) is no longer needed
        c.w_to_tau(w, out=Wqt)
        S = c.sigma(self.G, Wqt, self.X, hermitize=True)
        return S, st

    def step(self, lri=None, krylov=False, record=False):
        t0 = time.perf_counter()
        c = self.ctx
        if self.with_hf:
            self.F = self.H + c.hf(self.Dm, self.X, self.V)
            c.hermitize(self.F)
        if lri is None:
            S_new, st = self._sigma_direct()
        else:
            S_new, st = self._sigma_learned(lri, krylov, record)
        self.torch.cuda.synchronize()
        t_sigma = time.perf_counter() - t0
        dS = float((S_new - self.Sigma).abs().max())
        self.Sigma = S_new if self.mixing == 1.0 else (1.0 - self.mixing) * self.Sigma + self.mixing * S_new
        self.mu, nel = c.update_mu(self.F, self.Sigma, self.nel_target, self.mu)
        self.G, self.Dm = c.dyson_g(self.F, self.Sigma, self.mu)
        e_corr = c.gm_energy(self.G, self.Sigma)
        self.torch.cuda.synchronize()
        rec = dict(t_total=time.perf_counter() - t0, t_sigma=t_sigma, sigma_diff=dS, mu=self.mu, nelec=nel,
                   e_corr=e_corr, lri=st)
        self.history.append(rec)
        return rec

    def run(self, niter, dyson="direct", lri=None, train_iters=2, train_kw=None, mode="learned", max_iter=200,
            tol=1e-10, log=None):
        """niter iterations; with dyson="learned" an LearnedDyson is created if none is given.
        Returns the per-iteration records."""
        if dyson == "learned" and lri is None:
            lri = LearnedDyson(self.ctx)
        for it in range(1, niter + 1):
            if dyson == "direct":
                rec = self.step()
            else:
                krylov = it > train_iters
                if krylov and it == train_iters + 1 and mode == "learned":
                    t0 = time.perf_counter()
                    lf, ll = lri.train(**(train_kw or {}))
                    self.torch.cuda.synchronize()
                    self.train_info = dict(loss_first=lf, loss_last=ll, t_train=time.perf_counter() - t0,
                                           pairs=lri.num_pairs())
                if krylov:
                    lri.set(mode, max_iter=max_iter, tol=tol)
                rec = self.step(lri, krylov=krylov, record=not krylov)
            rec["it"] = it
            if log:
                log(rec)
        return self.history
