"""Learned form: the network's HIP kernels against the CPU oracle, the Krylov Dyson solve
against the direct solve and the FP64 oracle, its fallback, and what training gains.

Tolerances
  network forward and gradients: 2e-3 relative (Frobenius, per parameter block).  Both
    sides round the same operands to BF16; they differ by FP32 summation order and by the
    BF16 roundings that this moves across a rounding boundary.
  Adam: 1e-5 relative on the parameter change.
  Krylov W: 1e-8 relative to the largest element of the direct solve (every column's true
    residual is <= 1e-10 before a Krylov solve is accepted) and the same against the oracle.
"""
import numpy as np
import pytest

from coqui_amd import IAFT, LearnedDyson, Mlp, ScGW, ThcGW, make_system

pytestmark = pytest.mark.gpu

TABLE = "1e2.1e-15"


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


def frob_rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("dims,batch", [((70, 96, 40), 200), ((193, 512, 64), 300)])
def test_mlp_forward_backward_vs_oracle(torch, dims, batch):
    from oracle import mlp_oracle as mo

    print("case: mlp dims=%s batch=%d" % (dims, batch), flush=True)
    net = Mlp(*dims, seed=11)
    rng = np.random.default_rng(5)
    p = net.params().astype(np.float64)
    X = rng.standard_normal((batch, dims[0])).astype(np.float32)
    dY = rng.standard_normal((batch, dims[2])).astype(np.float32)
    Y = net.forward(X).cpu().numpy()
    g = net.backward(X, dY).cpu().numpy()
    Yr, cache = mo.forward(p, dims, X)
    gr = mo.backward(cache, dims, dY)
    print("forward err %.2e" % frob_rel(Y, Yr))
    assert frob_rel(Y, Yr) < 2e-3
    o, _ = mo.layout(*dims)
    keys = ["W0", "b0", "W1", "b1", "W2", "b2"]
    bounds = [o[k] for k in keys] + [len(gr)]
    for k, lo, hi in zip(keys, bounds[:-1], bounds[1:]):
        e = frob_rel(g[lo:hi], gr[lo:hi])
        print("grad %s err %.2e" % (k, e))
        assert e < 2e-3


def test_mlp_adam_vs_oracle(torch):
    from oracle import mlp_oracle as mo

    dims = (33, 64, 17)
    net = Mlp(*dims, seed=3)
    rng = np.random.default_rng(2)
    w = net.params().astype(np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    w0 = w.copy()
    for step in range(1, 4):
        g = (rng.standard_normal(w.size) * 0.1).astype(np.float32)
        net.adam_step(torch.from_numpy(g).cuda(), 1e-3)
        w, m, v = mo.adam(w, m, v, g.astype(np.float64), step, 1e-3)
    got = net.params().astype(np.float64)
    assert frob_rel(got - w0, w - w0) < 1e-5


def test_mlp_training_reduces_loss(torch):
    """Fit y = tanh(x A) with the fused train_mse step (forward, MSE gradient, backward, Adam)."""
    dims = (16, 128, 8)
    net = Mlp(*dims, seed=4)
    rng = np.random.default_rng(9)
    A = rng.standard_normal((16, 8)) / 4
    X = rng.standard_normal((512, 16)).astype(np.float32)
    T = np.tanh(X @ A).astype(np.float32)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    l0 = net.train_mse(Xd, Td, 3e-3)
    for _ in range(300):
        l1 = net.train_mse(Xd, Td, 3e-3)
    print("mse %.3e -> %.3e" % (l0, l1))
    assert l1 < 0.05 * l0


def small_problem(orc, torch, kgrid=(2, 2, 1), no=8, nmu=64, sid=61):
    s = make_system(kgrid, no, nmu, system_id=sid)
    tab = orc.load_ir(1000.0, None, table=TABLE)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    G = torch.from_numpy(orc.update_G(tab, s.C, s.E)).cuda()
    X = torch.from_numpy(s.X).cuda()
    V = torch.from_numpy(s.V).cuda()
    Piw = ctx.tau_to_w(ctx.pi(G, X))
    return s, ctx, Piw, V


def test_lri_krylov_matches_direct_and_oracle(orc, torch):
    print("case: lri krylov modes", flush=True)
    s, ctx, Piw, V = small_problem(orc, torch)
    ref = orc.dyson(Piw.cpu().numpy(), s.V)
    lri = LearnedDyson(ctx, rank=16, width=128)
    Wd = Piw.clone()
    lri.dyson(Wd, V, krylov=False, keep_state=True)
    Wd = Wd.cpu().numpy()
    assert relmax(Wd, ref) < 1e-10
    its = {}
    for mode in ("none", "woodbury", "previous", "learned"):
        lri.set(mode, max_iter=300, tol=1e-10)
        W = Piw.clone()
        st = lri.dyson(W, V, krylov=True, keep_state=True)
        W = W.cpu().numpy()
        its[mode] = st["iter_mean"]
        print(mode, st, "err vs direct %.2e" % relmax(W, Wd))
        assert st["fallbacks"] == 0 and st["krylov_accepted"] == st["solves"] == Piw.shape[0] * Piw.shape[1]
        assert st["resid_max"] <= 1e-10
        assert relmax(W, Wd) < 1e-8
        assert relmax(W, ref) < 1e-8
    assert its["woodbury"] <= its["none"]


def test_lri_fallback_to_direct(orc, torch):
    """One CG iteration cannot reach 1e-10: every chunk is re-solved directly."""
    s, ctx, Piw, V = small_problem(orc, torch, sid=62)
    lri = LearnedDyson(ctx, rank=16, width=64)
    lri.set("woodbury", max_iter=1, tol=1e-10)
    W = Piw.clone()
    st = lri.dyson(W, V, krylov=True)
    print(st)
    assert st["fallbacks"] == st["solves"] and st["krylov_accepted"] == 0
    assert relmax(W.cpu().numpy(), orc.dyson(Piw.cpu().numpy(), s.V)) < 1e-10


def test_learned_core_gain(orc, torch):
    """Self-consistent loop: iterations 1-2 direct (recording pairs), training, then at
    iteration 3 the core of every mode against the exact core, before and after training,
    and the CG iterations of each mode.  What training gains is asserted against the
    untrained network, the unpreconditioned solve and the Woodbury core (measured: core
    error 0.021 against 0.035, CG iterations 20 against 21); the carried-over core is
    printed."""
    print("case: learned gain scGW kgrid=(2,2,1) no=8 nmu=64", flush=True)
    kgrid, no, nmu = (2, 2, 1), 8, 64
    s = make_system(kgrid, no, nmu, system_id=63)
    ctx = ThcGW(kgrid, no, nmu, IAFT.from_table(1000.0, TABLE))
    lri = LearnedDyson(ctx, rank=16, width=512, seed=5)
    sc = ScGW(ctx, s)
    for it in (1, 2):
        sc.step(lri, krylov=False, record=True)
    assert lri.num_pairs() == 2 * ctx.nw_half * ctx.nk
    a = ctx.pi(sc.G, sc.X)
    Piw = ctx.tau_to_w(a)

    def core_err(mode):
        lri.set(mode)
        W = Piw.clone()
        return lri.dyson(W, sc.V, krylov=False, keep_state=True)["core_err"]

    def cg_iters(mode):
        lri.set(mode, max_iter=300, tol=1e-10)
        W = Piw.clone()
        st = lri.dyson(W, sc.V, krylov=True, keep_state=True)
        assert st["fallbacks"] == 0
        return st["iter_mean"]

    untrained = core_err("learned")
    lf, ll = lri.train(epochs=60, batch=32, lr=1e-3)
    errs = {m: core_err(m) for m in ("woodbury", "previous", "learned")}
    iters = {m: cg_iters(m) for m in ("none", "woodbury", "previous", "learned")}
    print("train loss %.3e -> %.3e" % (lf, ll))
    print("core error: untrained network %.3e, %s" % (untrained, errs))
    print("CG iterations:", iters)
    assert ll < 0.5 * lf
    assert errs["learned"] < untrained
    assert iters["learned"] <= iters["none"]
    assert errs["learned"] < 0.8 * errs["woodbury"]
    assert iters["learned"] <= iters["woodbury"]
