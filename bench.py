"""Benchmark: wall time of one THC G0W0 iteration on one MI355X.

Workload (default, --config k64-mu1600): N_k = 64 (4x4x4), N_orb = 104, N_mu = 1600,
N_tau = 80 (IR table lambda = 1e3, eps = 5e-18, built by coqui_amd/data/ir/gen_ir_tables.py),
beta = 1000, complex128 inputs and outputs.  One step = chi0 -> tau_to_w -> Dyson W (+ RPA
Ec from the same factors) -> w_to_tau -> Sigma (+ hermitize), with G(tau,k), X and V
resident in HBM.

Forms (--forms; --headline names the one reported as `value`; errors are taken against
the plain form, which runs first):
  plain      FP64 throughout, Dyson by LU with partial pivoting (the reference algorithm),
             complex products by four real MFMA products;
  plain_g3   the plain form with complex-FP64 products by three real MFMA products (flag 12,
             Gauss's form), as every form below uses (--g3-mask, default 13: also the FP32
             products with FP64 folding, i.e. the back-projection of Sigma, and Sigma's FP32
             projections of G);
  hermitian  FP64 throughout, Dyson by the two-Cholesky Hermitian form;
  mixed      FP32 products (projections, k-grid transforms, Dyson products), FP64 tau<->iW
             transforms, Cholesky factorisations and ln det;
  mixed_bf16 as mixed, with the FP32 products done as split-BF16 MFMA products;
  resolvent  FP64 throughout, Dyson by the resolvent form W + V = (V^-1 - Pi)^-1 (one
             Cholesky factorisation, triangular inverse and triangular product per matrix),
             and the projections of the Hermitian G on their lower triangle (flag 10);
  mixed_resolvent  FP32 projections and k-grid transforms, FP64 resolvent Dyson step;
  *_trev     the form with the Dyson step on the time-reversal irreducible q only (needs
             --trev, which draws the system time-reversal symmetric);
  *_c4       the form with the Dyson step on the q irreducible under a fourfold rotation
             (needs --c4, which draws the system symmetric under it);
  *_ibz      (plain_ibz, resolvent_ibz; needs --c4) the whole iteration on the irreducible
             wedge of the fourfold rotation: G and Sigma on the irreducible k, chi0, W and
             the Dyson step on the irreducible q (thcgw_gw_iteration_ibz); its errors are
             taken at the irreducible points (no QP energies).
Every form is reported with its speedup over plain and its errors against plain (W at
q = 0, Sigma, Ec, and the G0W0 quasiparticle energies in meV).

Prints one JSON line.  Also reported: the dominant GEMM class of the headline form measured
with HIP events on its launch stream during the timed steps (roofline), and the CPU
baseline: the reference algorithm on the host's OpenBLAS (oracle/blas_port.py) timed on a
bounded sample of the same workload and extrapolated to one iteration.
"""
import argparse
import faulthandler
import json
import os
import sys
import time

import numpy as np

faulthandler.enable()

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: kgrid, N_orb, N_mu, IR table (N_tau)
    "k8-mu200": ((2, 2, 2), 26, 200, "1e2.1e-15"),  # 40 tau points
    "k64-mu800": ((4, 4, 4), 52, 800, "1e3.5e-18"),  # 80 tau points
    "k64-mu1600": ((4, 4, 4), 104, 1600, "1e3.5e-18"),  # 80 tau points
    "batch24": None,  # batched: see batch_systems()
    "scgw-k64-mu1200": ((4, 4, 4), 78, 1200, "1e3.5e-18"),  # self-consistent GW, learned Dyson form: see run_scgw()
}
# The batched workload: 60 tau points (lambda = 1e3, eps = 1e-12).
BATCH_TABLE = "1e3.1e-12"


def batch_systems(n=24):
    """The batched workload (coqui_amd.system.batch_specs)."""
    from coqui_amd.system import batch_specs

    return batch_specs(n)


# form: (flag 6 Dyson algorithm, flag 7 precision, symmetry: 0 none, 1 time reversal (flag 9),
# 2 the C4 q-point map (thcgw_set_qsym), 3 the C4 irreducible wedge (thcgw_set_bz),
# flag 12: complex-FP64 products by 3 real MFMA products instead of 4)
FORMS = {"plain": (0, 0, 0, 0), "plain_g3": (0, 0, 0, 1), "hermitian": (1, 0, 0, 1), "mixed": (1, 1, 0, 1),
         "mixed_bf16": (1, 2, 0, 1), "resolvent": (2, 0, 0, 1), "mixed_resolvent": (2, 1, 0, 1),
         "plain_trev": (0, 0, 1, 0), "hermitian_trev": (1, 0, 1, 1), "mixed_trev": (1, 1, 1, 1),
         "resolvent_trev": (2, 0, 1, 1), "mixed_resolvent_trev": (2, 1, 1, 1),
         "plain_c4": (0, 0, 2, 0), "hermitian_c4": (1, 0, 2, 1), "mixed_c4": (1, 1, 2, 1),
         "resolvent_c4": (2, 0, 2, 1), "mixed_resolvent_c4": (2, 1, 2, 1),
         "plain_ibz": (0, 0, 3, 0), "resolvent_ibz": (2, 0, 3, 1)}
HA_PER_MEV = 1.0 / 27211.386245988
# GEMM profile class -> kernel template as rocprofv3 names it (64x64 tiles)
OPS = {"N": 0, "T": 1, "C": 2}


def gemm_kernel_name(cls, g3=False):
    if cls.endswith("32"):
        return "gemm_kernel<float, c64, c64, c64, %d, %d, 64, 64, false>" % (OPS[cls[0]], OPS[cls[1]])
    if cls.endswith("bf"):
        return "gemm_bf16x6_kernel<c64, c64, c64, %d, %d, false>" % (OPS[cls[0]], OPS[cls[1]])
    return "gemm_kernel<double, c128, c128, c128, %d, %d, 64, 64, false%s>" % (OPS[cls[0]], OPS[cls[1]],
                                                                             ", true" if g3 else "")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline_child(path):
    """Runs in a separate process that never touches the GPU (python bench.py --cpu-child F)."""
    from coqui_amd import IAFT

    z = np.load(path)
    kgrid = tuple(int(v) for v in z["kgrid"])
    ia = IAFT.from_table(float(z["beta"]), str(z["table"]))
    print(json.dumps(cpu_baseline(kgrid, z["X"], z["V2"], ia, z["G_pair"])), flush=True)


def run_cpu_baseline(s, ia, G_pair):
    """The CPU sample in a child process that never initialises the GPU, so that host BLAS
    threads do not share the process with the HIP runtime's."""
    import subprocess
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cpu_inputs.npz")
        np.savez(path, kgrid=np.array(s.kgrid), X=s.X, V2=s.V_host2, G_pair=G_pair, beta=ia.beta, table=ia.name)
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", path], capture_output=True,
                           text=True, timeout=900)
        if r.returncode != 0:
            log("bench: CPU baseline child failed (rc %d): %s" % (r.returncode, r.stderr[-2000:]))
            return None
        return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(kgrid, X, V2, ia, G_pair, reps=2):
    """The reference algorithm on the host's BLAS/LAPACK (oracle/blas_port.py: ZGEMM
    projections and dense k <-> R transforms, getrf + getri Dyson, getrf determinants,
    ZGEMM tau <-> iW) on a bounded sample, each part timed `reps` times (mean): one tau slice
    of chi0, one tau pair of Sigma, 2 Dyson and 2 Ec matrices, one q block of the
    imaginary-axis transforms; extrapolated to a full iteration."""
    import threadpoolctl

    from oracle import blas_port, oracle

    tab = oracle.load_ir(ia.beta, None, table=ia.name)
    Tp, Tm = blas_port.fold_tables(tab)
    nk = int(np.prod(kgrid))
    nmu = X.shape[1]
    nth, nwh = ia.nt_half, ia.nw_half
    rng = np.random.default_rng(0)
    t = {}
    sample_s = 0.0

    def timed(key, fn, count):
        nonlocal sample_s
        out = None
        t0 = time.perf_counter()
        for _ in range(count):
            out = fn()
        dt = time.perf_counter() - t0
        sample_s += dt
        t[key] = dt / count
        return out

    Pi1 = timed("pi_slice", lambda: blas_port.pi_pair(G_pair, X, kgrid), reps)
    timed("sigma_pair", lambda: blas_port.sigma_pair(G_pair, X, Pi1, kgrid), reps)  # any W of the right shape
    del Pi1
    B = rng.standard_normal((1, 2, nmu, nmu)) + 1j * rng.standard_normal((1, 2, nmu, nmu))
    Pi = -(B @ np.conj(np.swapaxes(B, -1, -2))) / nmu * 0.05
    timed("dyson_2", lambda: blas_port.dyson(Pi, V2), reps)
    timed("ec_2", lambda: blas_port.rpa_terms(Pi, V2), reps)
    Xt = rng.standard_normal((nth, nmu * nmu)) + 0j
    timed("ft_q", lambda: blas_port.w_to_tau(Tm, blas_port.tau_to_w(Tp, Xt)), reps)
    total = (t["pi_slice"] * nth + (t["dyson_2"] + t["ec_2"]) / 2.0 * nwh * nk + t["ft_q"] * nk
             + t["sigma_pair"] * nth)
    threads = max([p.get("num_threads", 1) for p in threadpoolctl.threadpool_info()
                   if p.get("internal_api") == "openblas"] or [1])
    return {
        "value": total,
        "unit": "s",
        "cores": int(threads),
        "kind": "port",
        "sample": ("reference algorithm on the host's OpenBLAS (numpy/scipy, %d threads): one tau slice of chi0, "
                   "one tau pair of Sigma, 2 Dyson + 2 Ec (q, iW) matrices and one q block of tau<->iW, each "
                   "timed %d times (%.1f s in all), extrapolated from the mean times to %d tau slices, %d (q, iW) "
                   "matrices and %d q blocks" % (threads, reps, sample_s, nth, nwh * nk, nk)),
        "sample_times_s": {k: round(v, 3) for k, v in t.items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="k64-mu1600", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", action="store_true", help="also time each stage of the headline form separately")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--gemm-tile", type=int, default=0, help="GEMM tile shape: 0 auto, 1 64x64, 2 128x64, 3 128x128")
    ap.add_argument("--gemm-tile32", type=int, default=0, help="FP32 GEMM tile shape: 0 64x64, 1 128x64, 2 128x128")
    ap.add_argument("--dyson-chunk", type=int, default=0, help="matrices per Dyson batch (0 = sized from free memory)")
    ap.add_argument("--pair-chunk", type=int, default=0, help="tau pairs per projection batch (0 = sized from free memory)")
    ap.add_argument("--dyson-lanes", type=int, default=1, choices=(1, 2),
                    help="resolvent Dyson chunks pipelined over this many streams (flag 11)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="after the timed steps, copy 1 GiB device to device (HBM counter calibration)")
    ap.add_argument("--scgw-iters", type=int, default=10, help="scgw-k64-mu1200: self-consistent iterations")
    ap.add_argument("--trev", action="store_true", help="draw the system time-reversal symmetric (for *_trev forms)")
    ap.add_argument("--c4", action="store_true", help="draw the system symmetric under a fourfold rotation (*_c4 forms)")
    ap.add_argument("--krylov-max-iter", type=int, default=60, help="scgw-k64-mu1200: CG iterations before the direct fallback")
    ap.add_argument("--compare-modes", type=int, default=1,
                    help="scgw-k64-mu1200: also solve the final state's Dyson step with every preconditioner (0 to skip)")
    ap.add_argument("--batch-lanes", type=int, default=4,
                    help="batch24: streams the systems are spread over (parallel branches of the graph)")
    ap.add_argument("--batch-dyson", default="resolvent", choices=("hermitian", "resolvent"),
                    help="batch24: Dyson form of the mixed-precision batch (FP32 Hermitian or FP64 resolvent)")
    ap.add_argument("--g3-mask", type=int, default=13,
                    help="flag 12 of the forms with three-product GEMMs: 1 FP64, | 2 FP32, | 4 FP32 with FP64 "
                         "folding, | 8 Sigma's FP32 projections (default 13: not the FP32 chi0 projections, where "
                         "three products triple the Ec error)")
    ap.add_argument("--batch-g3-mask", type=int, default=1,
                    help="batch24: flag 12 of the batched pass (default 1, FP64 only: with 5 the batch is no faster "
                         "and its largest QP error grows from 0.63 to 0.93 meV)")
    ap.add_argument("--forms", default="plain,plain_g3,mixed_resolvent",
                    help="comma-separated forms to time (plain, plain_g3, hermitian, mixed, mixed_bf16, resolvent, "
                         "mixed_resolvent, and their _trev / _c4 variants); plain runs first")
    ap.add_argument("--headline", default="mixed_resolvent", help="the form reported as value")
    args = ap.parse_args()
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    for f in forms:
        if f not in FORMS:
            ap.error("unknown form %r (choose from %s)" % (f, ", ".join(FORMS)))
        if FORMS[f][2] == 1 and not args.trev:
            ap.error("form %r needs --trev" % f)
        if FORMS[f][2] in (2, 3) and not args.c4:
            ap.error("form %r needs --c4" % f)
    if args.headline not in forms:
        forms.append(args.headline) if args.headline in FORMS else ap.error("unknown headline %r" % args.headline)
    if "plain" in forms:  # the reference of every error runs first
        forms = ["plain"] + [f for f in forms if f != "plain"]
    if args.cpu_child:
        cpu_baseline_child(args.cpu_child)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU sharding); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from coqui_amd import IAFT, ThcGW, make_system
    from coqui_amd.system import make_system_bz

    if args.config == "batch24":
        return run_batch(args)
    if args.config == "scgw-k64-mu1200":
        return run_scgw(args)
    kgrid, no, nmu, table = CONFIGS[args.config]
    beta = 1000.0
    log("bench: config %s kgrid=%s no=%d nmu=%d table=%s forms=%s headline=%s" % (
        args.config, kgrid, no, nmu, table, forms, args.headline))
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    bz = None
    if args.c4:
        s, bz = make_system_bz(kgrid, no, nmu, system_id=0, kind="c4")
        s.Z = None
    else:
        s = make_system(kgrid, no, nmu, system_id=0, with_v=False, trev=args.trev)
    ia = IAFT.from_table(beta, table)

    def new_ctx():
        c = ThcGW(kgrid, no, nmu, ia)
        c.set_flag(5, args.gemm_tile)
        c.set_flag(8, args.gemm_tile32)
        c.set_flag(3, args.dyson_chunk)
        c.set_flag(2, args.pair_chunk)
        c.set_flag(11, args.dyson_lanes)
        return c

    ctx = new_ctx()
    if s.V is not None:
        Vd = torch.from_numpy(s.V).cuda()
    else:
        Zd = torch.from_numpy(s.Z).cuda()
        Vd = torch.matmul(Zd, Zd.conj().transpose(1, 2)) / nmu  # input preparation, V = Z Z^H / N_mu
        del Zd
    Xd = torch.from_numpy(s.X).cuda()
    Gd = ctx.update_G(s.C, s.E, s.mu)
    s.V_host2 = Vd[:2].cpu().numpy()
    G_pair = Gd[[0, ia.nt - 1]].cpu().numpy()
    nth, nwh = ia.nt_half, ia.nw_half
    Sig = torch.empty((ia.nt, 1, s.nk, no, no), dtype=torch.complex128, device="cuda")
    Wqt = torch.empty((s.nk, nth, nmu, nmu), dtype=torch.complex128, device="cuda")
    H = torch.from_numpy(s.H).cuda()
    Cd = torch.from_numpy(s.C).cuda()
    if bz is not None:  # the irreducible wedge's inputs and outputs
        k_ibz = torch.from_numpy(bz["order"][:bz["nk_ibz"]]).cuda()
        q_ibz = torch.from_numpy(bz["order"][:bz["nq_ibz"]]).cuda()
        Gi = Gd[:, :, k_ibz].contiguous()
        Xl = Xd[torch.from_numpy(bz["order"]).cuda()].contiguous()
        Vi = Vd[q_ibz].contiguous()
        Sig_i = torch.empty((ia.nt, 1, bz["nk_ibz"], no, no), dtype=torch.complex128, device="cuda")
        W_i = torch.empty((bz["nq_ibz"], nth, nmu, nmu), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    log("bench: inputs ready in %.1f s" % (time.perf_counter() - t0))

    stream = torch.cuda.current_stream()

    def set_form(form):
        ctx.set_flag(6, FORMS[form][0])
        ctx.set_flag(7, FORMS[form][1])
        ctx.set_flag(10, int(FORMS[form][0] == 2))  # the resolvent forms also project G as Hermitian
        ctx.set_flag(9, int(FORMS[form][2] == 1))
        ctx.set_flag(12, FORMS[form][3] * args.g3_mask if FORMS[form][3] else 0)  # process-wide
        if FORMS[form][2] == 2:
            ctx.set_qsym(*s.qsym)
        else:
            ctx.set_qsym()
        if FORMS[form][2] == 3:
            ctx.set_bz(bz)

    def iterate(form):
        if FORMS[form][2] == 3:
            return ctx.gw_iteration_ibz(Gi, Xl, Vi, with_ec=True, hermitize=True, Sigma=Sig_i, W_qt=W_i)[2]
        return ctx.gw_iteration(Gd, Xd, Vd, with_ec=True, hermitize=True, Sigma=Sig, W_qt=Wqt)[2]

    def run_form(form, steps, warmup, first):
        # every form after the first gets a fresh context, so that no form runs with the
        # workspace of the forms before it held on the device (the Dyson chunk size follows
        # the free memory)
        nonlocal ctx
        if not first:
            ctx.close()
            torch.cuda.synchronize()
            ctx = new_ctx()
        set_form(form)
        ec = None
        for i in range(warmup):
            ec = iterate(form)
            log("bench: %s warmup %d done, Ec = %.12f" % (form, i, ec))
        ctx.profile_read()
        ctx.set_flag(4, 1)  # GEMM launch events during the timed steps
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for i in range(steps):
            ec = iterate(form)
            log("bench: %s step %d done" % (form, i))
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        ctx.set_flag(4, 0)
        return ev0.elapsed_time(ev1) / steps, wall, ec, ctx.profile_read()

    def classes(prof):
        return {k: {"ms_per_step": round(v[0] / args.steps, 2), "launches": v[2],
                    "tflops": round(v[1] / (v[0] * 1e-3) / 1e12, 2) if v[0] > 0 else 0.0}
                for k, v in prof.items()}

    results, ref = {}, None
    for i, form in enumerate(forms):
        f_ms, f_wall, f_ec, f_prof = run_form(form, args.steps, args.warmup, i == 0)
        wedge = FORMS[form][2] == 3
        E_f = None if wedge else ctx.solve_qp(Sig, Cd, H, s.E, mu=s.mu)
        r = {"ms_per_step": round(f_ms, 3), "wall_s_timed": round(f_wall, 3), "Ec_Ha": f_ec, "prof": f_prof,
             "gemm_classes": classes(f_prof)}
        if ref is None:
            ref = {"form": form, "ms": f_ms, "Sigma": Sig.clone(), "W0": Wqt[0].clone(), "ec": f_ec, "E": E_f}
        elif wedge:  # at the irreducible points (q = 0 is the first irreducible q)
            rel = lambda a, b: float((a - b).abs().max() / b.abs().max())
            r.update({
                "speedup_vs_%s" % ref["form"]: round(ref["ms"] / f_ms, 3),
                "err_W_q0_rel": rel(W_i[0], ref["W0"]),
                "err_Sigma_rel": rel(Sig_i, ref["Sigma"][:, :, k_ibz]),
                "err_Ec_Ha": abs(f_ec - ref["ec"]),
            })
        else:
            rel = lambda a, b: float((a - b).abs().max() / b.abs().max())
            r.update({
                "speedup_vs_%s" % ref["form"]: round(ref["ms"] / f_ms, 3),
                "err_W_q0_rel": rel(Wqt[0], ref["W0"]),
                "err_Sigma_rel": rel(Sig, ref["Sigma"]),
                "err_Ec_Ha": abs(f_ec - ref["ec"]),
                "err_qp_meV": float(np.max(np.abs(E_f - ref["E"])) / HA_PER_MEV),
                "err_qp_meV_median": float(np.median(np.abs(E_f - ref["E"])) / HA_PER_MEV),
                "qp_states_above_1meV": int(np.sum(np.abs(E_f - ref["E"]) > HA_PER_MEV)),
            })
        results[form] = r
        log("bench: %s %.1f ms per step %s" % (form, f_ms, json.dumps(
            {k: v for k, v in r.items() if k not in ("gemm_classes", "prof")})))
    del ref["Sigma"], ref["W0"]

    hd = results[args.headline]
    ms_step = hd["ms_per_step"]
    # roofline: the headline form's GEMM class with the most kernel time
    prof = hd["prof"]
    cls = max(prof, key=lambda k: prof[k][0])
    g_ms, g_fl, g_n, g_by = prof[cls]
    peak = FP32_PEAK_TFLOPS if cls.endswith("32") else FP64_PEAK_TFLOPS
    achieved = g_fl / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    a_ms = sum(v[0] for v in prof.values())
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "gemm_%s_%s_%s_traffic.json" % (cls, args.config, args.headline))
    if os.path.exists(tpath):  # HBM bytes per dispatch from the rocprofv3 FETCH_SIZE/WRITE_SIZE passes
        traffic = json.load(open(tpath))["traffic_bytes_per_dispatch"]
        traffic_src = os.path.relpath(tpath, ROOT)

    if args.copy_calib:
        xa = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        ya = torch.empty_like(xa)
        torch.cuda.synchronize()
        ya.copy_(xa)
        torch.cuda.synchronize()
        log("bench: copied 1 GiB (calibration)")
        del xa, ya

    stages = None
    if args.breakdown:
        # a fresh context, so that the buffers of the timed forms do not stay on the device
        ctx.close()
        torch.cuda.synchronize()
        ctx = new_ctx()
        set_form(args.headline)
        stages = {}
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        Pit = Wqt.view(nth, s.nk, nmu, nmu)  # the W_qt buffer doubles as Pi(tau), as in gw_iteration
        Piw = torch.empty((nwh, s.nk, nmu, nmu), dtype=torch.complex128, device="cuda")
        ev[0].record(stream)
        ctx.pi(Gd, Xd, out=Pit)
        ev[1].record(stream)
        ctx.tau_to_w(Pit, out=Piw)
        ev[2].record(stream)
        ctx.dyson_w(Piw, Vd, with_terms=True)
        ev[3].record(stream)
        ctx.w_to_tau(Piw, out=Wqt)
        ev[4].record(stream)
        ctx.sigma(Gd, Wqt, Xd, hermitize=True, out=Sig)
        ev[5].record(stream)
        torch.cuda.synchronize()
        for i, name in enumerate(["chi0", "tau_to_w", "dyson+ec_terms", "w_to_tau", "sigma"]):
            stages[name] = round(ev[i].elapsed_time(ev[i + 1]), 2)
        del Piw
        log("bench: stages (ms) %s" % stages)

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU baseline sample ...")
        cpu = run_cpu_baseline(s, ia, G_pair)
        if cpu:
            log("bench: CPU baseline %.1f s per iteration (extrapolated)" % cpu["value"])

    plain = results.get("plain")
    out = {
        "metric": "THC-GW iteration wall-time (s) + QP error (meV) vs FP64; speedup vs plain GPU",
        "value": round(ms_step / 1000.0, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c128" if FORMS[args.headline][1] == 0 else
                  "c128 in/out; FP32 projections and k-grid transforms, FP64 tau<->iW and Dyson"),
        "data": "synthetic (seeded generator, random H(k)/X/V; no checkpoint)",
        "config": {"workload": "%s %s THC G0W0 iteration (chi0, W, Sigma, RPA Ec)" % (args.config, args.headline),
                   "kgrid": list(kgrid), "nk": s.nk, "n_orb": no, "n_mu": nmu, "n_tau": ia.nt,
                   "n_iw_b": ia.nw_b, "ir_table": table, "beta": beta, "parallelism": "single GPU",
                   "time_reversal_symmetric_system": bool(args.trev), "c4_symmetric_system": bool(args.c4)},
        "form": args.headline,
        "Ec_Ha": hd["Ec_Ha"],
        "plain_s_per_iteration": round(plain["ms_per_step"] / 1000.0, 6) if plain else None,
        "speedup_vs_plain": (1.0 if args.headline == "plain" else hd.get("speedup_vs_plain")) if plain else None,
        "speedup_vs_plain_g3": (round(results["plain_g3"]["ms_per_step"] / ms_step, 3)
                                if "plain_g3" in results else None),
        "qp_error_meV": (0.0 if args.headline == "plain" else hd.get("err_qp_meV")) if plain else None,
        "ec_error_Ha": (0.0 if args.headline == "plain" else hd.get("err_Ec_Ha")) if plain else None,
        "forms": {k: {kk: vv for kk, vv in v.items() if kk != "prof"} for k, v in results.items()},
        "roofline": {"kernel": "%s (GEMM class %s of the %s form)" % (gemm_kernel_name(cls, bool(FORMS[args.headline][3] and args.g3_mask & 1)),
                                                                       cls, args.headline),
                     "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": round(g_by / max(g_n, 1)),
                     "flops_per_launch": round(g_fl / max(g_n, 1)),
                     "launches": g_n, "avg_launch_ms": round(g_ms / max(g_n, 1), 4),
                     "share_of_step_time": round(g_ms / max(ms_step * args.steps, 1e-9), 4),
                     "all_gemm_share_of_step_time": round(a_ms / max(ms_step * args.steps, 1e-9), 4)},
        "cpu_baseline": cpu,
        "stages_ms": stages,
        "wall_s_timed": hd["wall_s_timed"],
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def run_batch(args):
    """Batched form: the 24 systems of batch_systems() in mixed precision (FP32 projections
    and k-grid transforms; Dyson step by the FP64 resolvent form, or with --batch-dyson
    hermitian by the FP32 Hermitian form), one iteration of all of them per step, replayed
    from one captured HIP graph whose --batch-lanes streams (4) run as parallel branches;
    also timed one system after another without the graph (same chunk sizes), and
    against a plain FP64 pass of the same batch (LU Dyson, one system after another), which
    is also the reference of the errors."""
    import torch

    from coqui_amd import IAFT, ThcBatch, ThcGW, make_system

    torch.cuda.set_device(0)
    beta = 1000.0
    ia = IAFT.from_table(beta, BATCH_TABLE)
    specs = batch_systems()
    need = sum(int(np.prod(k)) * (ia.nt_half + 2) * nmu * nmu * 16 for k, no, nmu, _ in specs)
    log("bench: batch of %d systems, %.1f GB of per-system arrays" % (len(specs), need / 1e9))
    if need > 200e9:
        print("bench.py: the batch draw needs %.0f GB of per-system arrays" % (need / 1e9), file=sys.stderr)
        sys.exit(2)
    sysd = []
    t0 = time.perf_counter()
    for kgrid, no, nmu, sid in specs:
        s = make_system(kgrid, no, nmu, system_id=sid, with_v=False)
        ctx = ThcGW(kgrid, no, nmu, ia)
        Zd = torch.from_numpy(s.Z).cuda()
        Vd = torch.matmul(Zd, Zd.conj().transpose(1, 2)) / nmu
        del Zd
        d = dict(s=s, ctx=ctx, V=Vd, X=torch.from_numpy(s.X).cuda(), G=ctx.update_G(s.C, s.E, s.mu),
                 Sig=torch.empty((ia.nt, 1, s.nk, no, no), dtype=torch.complex128, device="cuda"),
                 W=torch.empty((s.nk, ia.nt_half, nmu, nmu), dtype=torch.complex128, device="cuda"))
        sysd.append(d)
    torch.cuda.synchronize()
    log("bench: batch inputs ready in %.1f s" % (time.perf_counter() - t0))
    lists = [[d[k] for d in sysd] for k in ("G", "X", "V", "Sig", "W")]
    b = ThcBatch([d["ctx"] for d in sysd], use_graph=True, tune=True)  # from here on the scratch is shared

    # plain FP64 reference pass, one system after another; then the same pass timed
    ref = []
    sysd[0]["ctx"].set_flag(12, 0)  # process-wide: four real products per complex product
    for d in sysd:
        d["ctx"].set_flag(6, 0)
        d["ctx"].set_flag(7, 0)
        S, W, ec = d["ctx"].gw_iteration(d["G"], d["X"], d["V"], with_ec=True, hermitize=True, Sigma=d["Sig"],
                                         W_qt=d["W"])
        H = torch.from_numpy(d["s"].H).cuda()
        Cd = torch.from_numpy(d["s"].C).cuda()
        E = d["ctx"].solve_qp(S, Cd, H, d["s"].E, mu=d["s"].mu)
        ref.append((S.clone(), W[0].clone(), ec, E))
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for d in sysd:
        d["ctx"].gw_iteration(d["G"], d["X"], d["V"], with_ec=True, hermitize=True, Sigma=d["Sig"], W_qt=d["W"])
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_plain = ev0.elapsed_time(ev1)
    log("bench: plain FP64 batch pass %.1f ms" % ms_plain)
    # the lanes (one scratch pool each, automatic chunks at 1/lanes of the memory) only for the
    # batched form: the plain pass above keeps the one-pool chunk sizes
    if args.batch_lanes != 1:
        b.set_lanes(args.batch_lanes)
    for d in sysd:
        d["ctx"].set_flag(6, 2 if args.batch_dyson == "resolvent" else 1)
        d["ctx"].set_flag(10, int(args.batch_dyson == "resolvent"))  # Hermitian G projections
        d["ctx"].set_flag(7, 1)  # mixed precision
    sysd[0]["ctx"].set_flag(12, args.batch_g3_mask)  # process-wide: complex products by three real products

    def timed(steps, graph):
        b_set = b.lib.thcgw_batch_set(b.h, int(graph), 1)
        assert b_set == 0
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ec = None
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(steps):
            ec = b.gw_iteration(*lists)
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / steps, ec

    t0 = time.perf_counter()
    b.gw_iteration(*lists)  # eager pass, tile tuning, capture and first replay
    log("bench: batch warm-up (tuning + capture) %.1f s" % (time.perf_counter() - t0))
    for _ in range(max(0, args.warmup - 1)):
        b.gw_iteration(*lists)
    ms_eager, _ = timed(max(1, args.steps), False)
    ms_graph, ec = timed(args.steps, True)
    errs = {"W_q0_rel": 0.0, "Sigma_rel": 0.0, "Ec_Ha": 0.0, "qp_meV": 0.0}
    dqp = []
    for d, (S0, W00, ec0, E0), e in zip(sysd, ref, ec):
        H = torch.from_numpy(d["s"].H).cuda()
        Cd = torch.from_numpy(d["s"].C).cuda()
        E = d["ctx"].solve_qp(d["Sig"], Cd, H, d["s"].E, mu=d["s"].mu)
        rel = lambda a, c: float((a - c).abs().max() / c.abs().max())
        errs["W_q0_rel"] = max(errs["W_q0_rel"], rel(d["W"][0], W00))
        errs["Sigma_rel"] = max(errs["Sigma_rel"], rel(d["Sig"], S0))
        errs["Ec_Ha"] = max(errs["Ec_Ha"], abs(e - ec0))
        errs["qp_meV"] = max(errs["qp_meV"], float(np.max(np.abs(E - E0)) / HA_PER_MEV))
        dqp.append(np.abs(E - E0).ravel() / HA_PER_MEV)
    dqp = np.concatenate(dqp)
    errs["qp_meV_median"] = float(np.median(dqp))
    errs["qp_states_above_1meV"] = int(np.sum(dqp > 1.0))
    errs["qp_states"] = int(dqp.size)
    out = {
        "metric": "THC-GW iteration wall-time (s) + QP error (meV) vs FP64; speedup vs plain GPU",
        "value": round(ms_graph / 1000.0, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_graph, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c128 in/out, FP32 projections and k-grid transforms, FP64 IR transforms and resolvent Dyson"
                  if args.batch_dyson == "resolvent" else
                  "c128 in/out, FP32 products, FP64 IR transforms and factorisations"),
        "data": "synthetic (seeded generator, random H(k)/X/V; no checkpoint)",
        "config": {"workload": "batched mixed-precision G0W0 iteration of %d systems (%s Dyson)" % (
                       len(specs), args.batch_dyson),
                   "systems": [{"kgrid": list(k), "n_orb": no, "n_mu": nmu} for k, no, nmu, _ in specs],
                   "ir_table": BATCH_TABLE, "n_tau": ia.nt, "beta": beta, "parallelism": "single GPU"},
        "batch_lanes": args.batch_lanes,
        "three_real_products_mask": args.batch_g3_mask,
        "ms_per_step_one_by_one": round(ms_eager, 3),
        "graph_speedup": round(ms_eager / ms_graph, 3),
        "ms_plain_fp64_one_by_one": round(ms_plain, 3),
        "speedup_vs_plain": round(ms_plain / ms_graph, 3),
        "max_errors_vs_plain_fp64": errs,
    }
    print(json.dumps(out), flush=True)
    b.close()
    for d in sysd:
        d["ctx"].close()


def run_scgw(args):
    """Learned form: self-consistent GW (scf_driver.cpp:96-170 order) for --scgw-iters
    iterations, first with the plain FP64 Dyson solve (LU, the reference algorithm), then
    with the learned form: direct Hermitian solves that record training pairs in iterations
    1-2, training of the network, and from iteration 3 preconditioned CG with the learned
    rank-64 approximate inverse, each solve accepted at a true FP64 residual <= 1e-10 or
    redone directly.  Reports the wall time per iteration of both, the CG iterations and
    fallbacks, and the difference of the final state (QP energies from the final Sigma,
    Galitskii-Migdal energy, chemical potential)."""
    import torch

    from coqui_amd import IAFT, LearnedDyson, ScGW, ThcGW, make_system

    kgrid, no, nmu, table = CONFIGS["scgw-k64-mu1200"]
    beta = 1000.0
    torch.cuda.set_device(0)
    ia = IAFT.from_table(beta, table)
    s = make_system(kgrid, no, nmu, system_id=0)
    log("bench: scGW kgrid=%s no=%d nmu=%d nt=%d, %d iterations" % (kgrid, no, nmu, ia.nt, args.scgw_iters))

    def final_state(ctx, sc):
        H = torch.from_numpy(s.H).cuda()
        Cd = torch.from_numpy(s.C).cuda()
        E = ctx.solve_qp(sc.Sigma, Cd, H, s.E, mu=sc.mu)
        return dict(E=E, mu=sc.mu, e_corr=sc.history[-1]["e_corr"], Sigma=sc.Sigma.clone())

    def show(tag):
        return lambda r: log("bench: %s it %d  %.2f s (Sigma %.2f s)  dSigma %.2e  mu %.6f  Ec(GM) %.8f  %s" % (
            tag, r["it"], r["t_total"], r["t_sigma"], r["sigma_diff"], r["mu"], r["e_corr"],
            "" if not r["lri"] else "CG %(iter_mean).1f fallbacks %(fallbacks)d accepted %(krylov_accepted)d" % r["lri"]))

    # plain FP64 GPU path
    ctx = ThcGW(kgrid, no, nmu, ia)
    sc = ScGW(ctx, s)
    hist_p = sc.run(args.scgw_iters, dyson="direct", log=show("plain"))
    ref = final_state(ctx, sc)
    del sc
    ctx.close()
    torch.cuda.empty_cache()

    # learned form
    ctx = ThcGW(kgrid, no, nmu, ia)
    sc = ScGW(ctx, s)
    lri = LearnedDyson(ctx, rank=64, width=512, seed=20261202)
    hist_l = sc.run(args.scgw_iters, dyson="learned", lri=lri, train_iters=2,
                    train_kw=dict(epochs=30, batch=256, lr=1e-3), mode="learned",
                    max_iter=args.krylov_max_iter, tol=1e-10, log=show("learned"))
    got = final_state(ctx, sc)
    kry = [r for r in hist_l if r["lri"] and r["lri"]["krylov_accepted"] + r["lri"]["fallbacks"] > 0]
    # one more Dyson step at the final state with every preconditioner and the direct solve
    modes = {}
    if args.compare_modes:
        a, w = sc._buffers()
        ctx.pi(sc.G, sc.X, out=a)
        ctx.tau_to_w(a, out=w)
        # the copy of Pi(iW), as large as w, is kept in host memory: the device holds w, the
        # library's Dyson chunk buffers (up to 60% of the free memory) and the training pairs
        Pi0 = w.cpu()
        for mode in ("direct", "none", "woodbury", "previous", "learned"):
            w.copy_(Pi0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if mode == "direct":
                st = lri.dyson(w, sc.V, krylov=False, keep_state=True)
            else:
                lri.set(mode, max_iter=args.krylov_max_iter, tol=1e-10)
                st = lri.dyson(w, sc.V, krylov=True, keep_state=True)
            torch.cuda.synchronize()
            modes[mode] = {"s": round(time.perf_counter() - t0, 3), "cg_iterations_mean": round(st["iter_mean"], 2),
                           "fallbacks": st["fallbacks"]}
            log("bench: final-state Dyson, %s: %s" % (mode, modes[mode]))
        del Pi0
    t_plain = sum(r["t_total"] for r in hist_p)
    t_learned = sum(r["t_total"] for r in hist_l)
    dqp = np.abs(got["E"] - ref["E"]) / HA_PER_MEV
    out = {
        "metric": "THC-GW iteration wall-time (s) + QP error (meV) vs FP64; speedup vs plain GPU",
        "value": round(t_learned / len(hist_l), 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": len(hist_l),
        "warmup": 0,
        "ms_per_step": round(1000.0 * t_learned / len(hist_l), 1),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (FP64 Krylov and direct solves); network BF16 weights, FP32 accumulation",
        "data": "synthetic (seeded generator, random H(k)/X/V; no checkpoint)",
        "config": {"workload": "self-consistent GW (10 iterations), learned Krylov Dyson form", "kgrid": list(kgrid),
                   "n_orb": no, "n_mu": nmu, "ir_table": table, "n_tau": ia.nt, "beta": beta,
                   "iterations": args.scgw_iters, "rank": 64, "width": 512, "parallelism": "single GPU"},
        "plain_s_per_iteration": round(t_plain / len(hist_p), 4),
        "speedup_vs_plain": round(t_plain / t_learned, 3),
        "learned_krylov_s_per_iteration": round(sum(r["t_total"] for r in kry) / max(1, len(kry)), 4),
        "cg_iterations_mean": [round(r["lri"]["iter_mean"], 1) for r in kry],
        "fallbacks": sum(r["lri"]["fallbacks"] for r in kry),
        "krylov_accepted": sum(r["lri"]["krylov_accepted"] for r in kry),
        "resid_max": max([r["lri"]["resid_max"] for r in kry] or [0.0]),
        "training": getattr(sc, "train_info", None),
        "final_state_dyson_by_mode": modes,
        "errors_vs_plain": {"qp_meV_max": float(dqp.max()), "qp_meV_median": float(np.median(dqp)),
                            "e_corr_gm_Ha": abs(got["e_corr"] - ref["e_corr"]), "mu_Ha": abs(got["mu"] - ref["mu"]),
                            "Sigma_rel": float((got["Sigma"] - ref["Sigma"]).abs().max() / ref["Sigma"].abs().max())},
    }
    print(json.dumps(out), flush=True)
    lri.close()
    ctx.close()


if __name__ == "__main__":
    main()
