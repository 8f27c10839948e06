#!/usr/bin/env python
"""Benchmark of the VMC inner loop on one MI355X.

Default workload (hubbard8x8_slater): 8x8 open-boundary Fermi-Hubbard, t=1,
U=8, 32 up + 32 down; Slater determinant with a 128x64 FP64 orbital matrix
drawn N(0, 0.1) from seed 0; 8192 walkers in one launch; 200 burn-in sweeps;
Sherman-Morrison updates with an exact FP64 refresh every 64 accepted moves;
SR in sample space with diagonal shift 1e-4, learning rate 0.05.

hubbard6x6_nnbf: 6x6, 16 up + 16 down; neural-network backflow 72 -> 256 ->
256 (ReLU) -> 72x32 correction (bf_scale 0.01) on an N(0, 0.1) base,
Xavier weights from seed 0, 678,912 parameters; BF16 GEMMs with FP32
accumulation and an FP64 determinant; 4096 walkers; O in FP32 (centred in
place); SR with a BF16 Gram preconditioning conjugate gradients on the exact
operator, shift 1e-3.

One timed step = one SR iteration: one sweep of every walker (one sample per
walker), the local energies, the per-sample log-derivatives O, the minSR solve
and the SGD update, all on the device.

hubbard_mixed_lp: the batched learned-proposal form.  An uneven batch of 4x4
(n = 1, 8192 walkers), 6x6 (n = 0.89, 4096) and 8x8 (n = 1, 4096) Hubbard
lattices, each with an NNBF ansatz of the 6x6 family (2N -> 256 -> 256 ->
2N x nf, BF16 GEMMs); one proposal net (three 3x3 convolutions, 32 channels)
serves all three and is trained on the chains' own walkers.  A step there is
one sample per walker of every lattice: |E|/4 learned proposals (the plain
form: one sweep of |E| edge proposals) and the local energies, with a
training step of the net every 10 steps; the wavefunction parameters stay
fixed (sampling form).  Reported for both samplers: tau_int, acceptance,
E/site and effective samples/s.

Printed: one JSON line with effective samples/s as `value`, SR iterations/s,
walker-sweeps/s, the roofline of the dominant component (HIP-event timing on
the launch stream), per-component timings, the CPU baseline (the FP64 oracle
restating the reference algorithm, timed on this host) and the |dE/site| of
the GPU local energies against the FP64 oracle.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12          # B/s
PEAK_FP64 = 78.6e12        # FLOP/s (vector = matrix on MI355X)
PEAK_FP32_MFMA = 157.3e12
PEAK_BF16 = 2.5e15

WORKLOADS = {
    "hubbard8x8_slater": dict(
        kind="slater", Lx=8, Ly=8, nu=32, nd=32, B=8192, burn=200, shift=1e-4, lr=0.05,
        gram="fp64",
        desc="8x8 OBC Hubbard t=1 U=8 32u+32d, Slater 128x64 N(0,0.1), 8192 walkers"),
    "hubbard4x2_slater": dict(
        kind="slater", Lx=4, Ly=2, nu=4, nd=4, B=1024, burn=100, shift=1e-4, lr=0.05,
        gram="fp64",
        desc="4x2 OBC Hubbard t=1 U=8 4u+4d, Slater 16x8 N(0,0.1), 1024 walkers"),
    "hubbard6x6_nnbf": dict(
        kind="nnbf", Lx=6, Ly=6, nu=16, nd=16, B=4096, burn=200, shift=1e-3, lr=0.05,
        gram="factored", hidden=256, bf_scale=0.01, mlp="bf16",
        desc=("6x6 OBC Hubbard t=1 U=8 16u+16d, NNBF 72-256-256-2304 ReLU, bf_scale 0.01 "
              "(678,912 params), BF16 GEMMs, FP64 Gram from the factors of O, 4096 walkers")),
    "hubbard8x8_nnbf": dict(
        kind="nnbf", Lx=8, Ly=8, nu=32, nd=32, B=4096, burn=200, shift=1e-3, lr=0.05,
        gram="factored", o_dtype="fp64", hidden=256, bf_scale=0.01, mlp="bf16",
        desc=("8x8 OBC Hubbard t=1 U=8 32u+32d, NNBF 128-256-256-8192 ReLU, bf_scale 0.01 "
              "(2,212,352 params), BF16 GEMMs, FP64 Gram from the factors of O, 4096 "
              "walkers")),
    "hubbard4x4_nnftns": dict(
        kind="peps", Lx=4, Ly=4, nu=8, nd=8, B=2048, burn=100, shift=5e-4, lr=0.1,
        gram="factored", D=4, hidden=128, eta=1.0,
        desc=("4x4 OBC Hubbard t=1 U=8 8u+8d, Z2 fPEPS D=4 (exact contraction, chi=16) with "
              "NN-fTNS backflow 32-128-128-3200 ReLU (436,736 params), FP64 (Gram from the "
              "factors of O), 2048 walkers")),
    "hubbard4x4_fpeps": dict(
        kind="peps", Lx=4, Ly=4, nu=8, nd=8, B=2048, burn=100, shift=5e-4, lr=0.1,
        gram="fp64", D=4, hidden=0, eta=0.0,
        desc=("4x4 OBC Hubbard t=1 U=8 8u+8d, Z2 fPEPS D=4 (exact contraction, chi=16), "
              "3200 params, FP64, 2048 walkers")),
    "hubbard6x6_fpeps_bmps": dict(
        kind="peps", Lx=6, Ly=6, nu=14, nd=14, B=1024, burn=20, shift=5e-4, lr=0.1,
        gram="fp64", D=4, chi=8, hidden=0, eta=0.0,
        desc=("6x6 OBC Hubbard t=1 U=8 14u+14d (N-8 electrons, as the reference's fPEPS "
              "script), Z2 fPEPS D=4, boundary MPS truncated to chi=8, FP64, 1024 walkers")),
    "hubbard_mixed_lp": dict(
        kind="mixed_lp", instances=[(4, 4, 8, 8, 8192), (6, 6, 16, 16, 4096), (8, 8, 32, 32, 4096)],
        hidden=256, bf_scale=0.01, mlp="bf16", burn=50, channels=32, train_iters=100,
        train_every=10, train_walkers=1024, train_lr=0.05,
        desc=("uneven batch of OBC Hubbard t=1 U=8 lattices: 4x4 8u+8d (8192 walkers), 6x6 "
              "16u+16d (4096), 8x8 32u+32d (4096); NNBF 2N-256-256-(2N x nf) ReLU, bf_scale "
              "0.01, BF16 GEMMs, per lattice; one learned proposal net (3 x 3x3 conv, 32 "
              "channels, BF16 matrix cores) shared by the three")),
}


# measured HBM traffic (bytes per call) of a roofline kernel:
# profiles/pmc_traffic_sr_gram_hubbard8x8_slater.md (Gram: SYRK + partial-sum pass)
TRAFFIC = {("hubbard8x8_slater", "sr.gram", "fp64"): 20.33e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="hubbard8x8_slater", choices=sorted(WORKLOADS))
    p.add_argument("--gram", default=None,
                   choices=["fp64", "fp32", "bf16", "bf16x3", "factored"],
                   help="Gram precision (default: the workload's); factored (NNBF): the "
                        "FP64 Gram from the Grams of the factors of O, no O matrix")
    p.add_argument("--refresh-every", type=int, default=64)
    p.add_argument("--chunk", type=int, default=65536,
                   help="NNBF: amplitudes per batched forward in the local energy")
    p.add_argument("--proposals-per-sample", type=int, default=0,
                   help="NNBF --sampler learned: proposals per sample (0 = |E| / 4)")
    p.add_argument("--lp-moves", type=int, default=1, choices=[1, 2],
                   help="NNBF --sampler learned: moves chained per proposal")
    p.add_argument("--sweeps-per-sample", type=int, default=1,
                   help="Slater: Metropolis sweeps between samples (1 = the reference's "
                        "one sample per sweep)")
    p.add_argument("--mlp", default=None, choices=["fp64", "fp32", "bf16"],
                   help="NNBF network precision (default: the workload's)")
    p.add_argument("--det", default=None, choices=["fp64", "fp32"],
                   help="NNBF determinant arithmetic of the amplitude-only passes (fp32: "
                        "BF16 MLP on the matrix-core layers only; default fp64)")
    p.add_argument("--slater-precision", default="fp64", choices=["fp64", "fp32"],
                   help="Slater walker state and sweep arithmetic (fp32: the FP32 form, "
                        "FP64 refresh and SR)")
    p.add_argument("--no-target", action="store_true",
                   help="skip the 8x8 NNBF target comparison of the default run")
    p.add_argument("--target-steps", type=int, default=8,
                   help="timed SR iterations per form in the target comparison")
    p.add_argument("--o-dtype", default=None, choices=["fp32", "fp64"],
                   help="NNBF log-derivative matrix O (default: the workload's, else fp32)")
    p.add_argument("--sampler", default="plain", choices=["plain", "learned"],
                   help="NNBF: plain sweeps, or the learned proposal (|E|/4 proposals per "
                        "sample, the net trained before and every 10 SR steps)")
    p.add_argument("--peps-reuse", action="store_true",
                   help="fPEPS: sweeps and local energies from the walkers' cached column "
                        "environments (fPEPS_Model_reuse_GPU)")
    p.add_argument("--peps-amp", default="fp64", choices=["fp64", "fp32"],
                   help="fPEPS / NN-fTNS: amplitudes contracted in FP64, or in FP32 on the "
                        "matrix cores (amp='fp32'; the log-derivative stays FP64)")
    p.add_argument("--tau-sweeps", type=int, default=64)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--copy-probe-mib", type=int, default=0,
                   help="first copy a buffer of this many MiB on the device (a plain copy "
                        "kernel of known size, to scale hardware byte counters)")
    return p.parse_args()


def log(msg):
    """Progress line on stderr (stdout carries only the JSON result)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def library_factor_times(O, shift, reps=2):
    """rocSOLVER (through torch.linalg) on the same SR matrix T = Oc Oc^T / Ns
    + shift I that the library factorises (Oc: O with its column means
    removed), timed after the measured loop for comparison with the
    library's blocked Cholesky: potrf (torch.linalg.cholesky) and the LU
    solve the reference calls (torch.linalg.solve, GPU/optimizer.py:546-563)."""
    import torch
    Ns = O.shape[0]
    Oc = O - O.mean(0)
    T = Oc @ Oc.T
    del Oc
    T /= Ns
    T.diagonal().add_(shift)
    e = torch.ones(Ns, dtype=torch.float64, device=O.device)
    out = {}
    for name, fn in (("torch_linalg_cholesky_ms", lambda: torch.linalg.cholesky(T)),
                     ("torch_linalg_solve_ms", lambda: torch.linalg.solve(T, e))):
        fn()
        torch.cuda.synchronize()
        s, t = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        t.record()
        torch.cuda.synchronize()
        out[name] = s.elapsed_time(t) / reps
    log("rocSOLVER factorisation of the same T timed: " +
        ", ".join(f"{k} {v:.2f}" for k, v in out.items()))
    return out


# ------------------------------------------------------------- CPU baseline
def _cpu_worker(args):
    """Reference algorithm on one core: batched full-amplitude sweep, E_loc by
    get_conn + full amplitudes, O_k, for a chunk of walkers."""
    os.environ["OMP_NUM_THREADS"] = "1"
    import numpy as np

    from oracle import hubbard as ohub
    from oracle import nnbf as onn
    from oracle import sampler as osamp
    from oracle import slater as osl
    kind, Lx, Ly, nu, nd, params, bf, x, seed, budget = args[:10]
    grad = args[10] if len(args) > 10 else True
    H = ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd)
    edges = H.graph.sweep_edges()
    if kind == "slater":
        M = params[0]

        def logamp(c):
            return osl.slogdet(M, c)

        def work(xs, s, l):
            osl.local_energy(H, M, xs, logamp=(s, l))
            osl.log_derivative(M, xs)
    elif kind == "nnbf":
        def logamp(c):
            return onn.forward(params, c, 2, bf)[:2]

        def work(xs, s, l):
            onn.local_energy(H, params, xs, 2, bf)
            if grad:
                onn.log_derivative(params, xs, 2, bf)
    else:
        from oracle import fpeps as ofp
        sh = ofp.PEPSShape(Lx, Ly, bf["D"])
        th, nn, eta = params[0], (params[1:] or None), bf["eta"]

        chi = bf.get("chi", 0)

        def amp(c):
            if chi > 0:
                t = th if nn is None else ofp.theta_eff(th, ofp.mlp_forward(nn, c)[0][0], eta)
                s_, l_ = ofp.bmps_log_amplitude(sh, t, c, chi)
                return s_ * np.exp(l_)
            return (ofp.amplitude(sh, th, c) if nn is None
                    else ofp.nnftn_amplitude(sh, th, nn, eta, c))

        def logamp(cs):
            a = np.array([amp(c) for c in np.atleast_2d(cs)])
            with np.errstate(divide="ignore"):
                return np.sign(a), np.maximum(np.log(np.abs(a)), np.log(1e-45))

        def work(xs, s, l):
            ofp.local_energy(H, sh, th, xs, nn, eta, chi=chi)
            for c in xs:
                if nn is None:
                    if chi > 0:
                        ofp.bmps_cheap_log_derivative(sh, th, c, chi)
                    else:
                        ofp.log_derivative(sh, th, c)
                else:
                    ofp.nnftn_log_derivative(sh, th, nn, eta, c, chi=chi)
    done, t0 = 0, time.perf_counter()
    chunk = 4
    while time.perf_counter() - t0 < budget:
        xs = x[(done % len(x)):(done % len(x)) + chunk]
        xs, s, l, _ = osamp.sweep(logamp, xs, edges, osamp.PhiloxDraws(seed, done), 1)
        work(xs, s, l)
        done += xs.shape[0]
    return done, time.perf_counter() - t0


def cpu_baseline(kind, Lx, Ly, nu, nd, params, bf, x, n_samples_step, tau, shift, seconds):
    import multiprocessing as mp

    import numpy as np

    from oracle import sr as osr
    # every core of this process's CPU share: one GPU's share of the box is
    # 16 cores (the processes beyond it would only time-slice)
    cores = min(len(os.sched_getaffinity(0)), 16)
    ctx = mp.get_context("spawn")
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    per = max(1, len(x) // cores)
    jobs = [(kind, Lx, Ly, nu, nd, params, bf, x[i * per:(i + 1) * per], 1000 + i,
             seconds * 0.6) for i in range(cores)]
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, jobs)
    samples = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    samples_per_s = samples / wall
    # SR on the host (numpy, multithreaded BLAS): the Gram O O^T timed at a
    # reduced sample count and scaled by Ns^2 (its cost is Ns^2 Np), the
    # dense solve of the Ns x Ns system (the reference's linalg.solve, LU,
    # 2/3 Ns^3) timed at the full Ns
    Np = int(sum(p.size for p in params))
    ns_cpu = min(2048, int(n_samples_step)) if Np < 100000 else 128
    rng = np.random.default_rng(0)
    O = rng.normal(size=(ns_cpu, Np))
    t0 = time.perf_counter()
    Oc = O - O.mean(axis=0)
    _ = Oc @ Oc.T
    t_gram_small = time.perf_counter() - t0
    del O, Oc
    Ns = int(n_samples_step)
    X = rng.normal(size=(Ns, 64))
    T = X @ X.T / 64 + shift * np.eye(Ns) + np.eye(Ns)
    eps = rng.normal(size=Ns)
    t0 = time.perf_counter()
    np.linalg.solve(T, eps)
    t_solve = time.perf_counter() - t0
    del X, T
    t_sr = t_gram_small * (n_samples_step / ns_cpu) ** 2 + t_solve
    t_iter = n_samples_step / samples_per_s + t_sr
    return {
        "value": n_samples_step / t_iter / tau,
        "unit": "effective samples/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"FP64 oracle restating the reference algorithm (full amplitude per "
                   f"proposal and per connected configuration), {samples} walker-samples "
                   f"(sweep + E_loc + O_k) on {cores} processes in {wall:.1f} s; minSR: "
                   f"Gram timed at Ns={ns_cpu}, Np={Np} ({t_gram_small:.2f} s) and scaled "
                   f"by Ns^2 to Ns={Ns}, plus the dense solve timed at Ns={Ns} "
                   f"({t_solve:.2f} s); same tau_int as the GPU chain"),
        "walker_sweeps_per_s": samples_per_s,
        "sr_iters_per_s": 1.0 / t_iter,
        "cpu_model": _cpu_model(),
    }


def _cpu_model():
    """The host CPU as /proc/cpuinfo names it (model name, family / model /
    stepping numbers) and the core counts (this process's share of all)."""
    f = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = line.split(":", 1)
                f.setdefault(k.strip(), v.strip())
    except OSError:
        pass
    name = f.get("model name", "unknown")
    ids = ", ".join(f"{k} {f[k]}" for k in ("vendor_id", "cpu family", "model", "stepping")
                    if k in f)
    return (f"{name} ({ids}); {len(os.sched_getaffinity(0))} of {os.cpu_count()} host "
            "cores in this process's share")


# ------------------------------------------------------------- workloads
class SlaterRun:
    """Slater walkers driven through the kernel wrappers (the C-ABI calls the
    reference-shaped model class makes, without its parameter-change checks)."""

    names = ["refresh", "sweep", "local_energy", "logderiv", "sr", "update"]

    def __init__(self, w, a, dev, H, fxs):
        import torch

        from vmc_torch_amd import kernels
        self.k, self.a, self.w, self.H, self.fxs = kernels, a, w, H, fxs
        N, nf, B = w["Lx"] * w["Ly"], w["nu"] + w["nd"], w["B"]
        self.N, self.nf, self.B, self.Np = N, nf, B, 2 * N * nf
        g = torch.Generator().manual_seed(0)
        self.M = (torch.randn(2 * N, nf, generator=g, dtype=torch.float64) * 0.1).to(dev)
        # the FP32 form: M rounded to FP32 (the model's parameters), A^-1 in FP32;
        # the exact refresh takes M widened back to FP64
        self.fp32 = getattr(a, "slater_precision", "fp64") == "fp32"
        if self.fp32:
            self.M32 = self.M.to(torch.float32)
            self.M = self.M32.to(torch.float64)
        self.e_sweep = H.graph.edge_tensor(dev, "sweep")
        self.e_graph = H.graph.edge_tensor(dev, "graph")
        self.st = kernels.SlaterWalkerState.empty(
            B, N, nf, dev, dtype=torch.float32 if self.fp32 else torch.float64)
        self.O = torch.empty(B, self.Np, dtype=torch.float64, device=dev)
        self.dp = torch.empty(self.Np, dtype=torch.float64, device=dev)
        self.ws = kernels.SRWorkspace()
        self.spp = max(1, a.sweeps_per_sample)
        self.proposals_per_step = self.e_sweep.shape[0] * self.spp

    def _sweep(self, n, seed, off):
        self.k.slater_sweep(self.M, self.e_sweep, self.fxs, self.st, n_sweeps=n, seed=seed,
                            sweep_offset=off, refresh_every=self.a.refresh_every,
                            M32=self.M32 if self.fp32 else None)

    def burn(self, n, seed, off):
        self.k.slater_refresh(self.M, self.fxs, self.st)
        self._sweep(n, seed, off)

    def local_energy(self):
        return self.k.slater_local_energy(self.M, self.e_graph, 1.0, 8.0, self.fxs, self.st,
                                          M32=self.M32 if self.fp32 else None)

    def step(self, mark, seed, off, gram, phases):
        m = [mark()]
        self.k.slater_refresh(self.M, self.fxs, self.st)     # M changed in the last update
        m.append(mark())
        acc0 = self.st.n_accept.clone()
        self._sweep(self.spp, seed, off)
        m.append(mark())
        E = self.local_energy()
        m.append(mark())
        self.k.slater_logderiv(self.st, self.N, self.nf, out=self.O)
        m.append(mark())
        _, stats, info = self.k.sr_minsr(self.O, E, None, self.w["shift"], gram,
                                         0 if gram == "fp64" else 20, 1e-8, self.ws,
                                         dp=self.dp, phase_ms=phases)
        m.append(mark())
        if self.fp32:
            self.k.sgd_update(self.M32.view(-1), self.dp, self.w["lr"])
            self.M.copy_(self.M32)
        else:
            self.k.sgd_update(self.M.view(-1), self.dp, self.w["lr"])
        m.append(mark())
        return m, E, self.st.n_accept - acc0, stats, info

    def pre_tau(self):
        self.k.slater_refresh(self.M, self.fxs, self.st)

    def tau_sample(self, seed, off):
        self._sweep(self.spp, seed, off)
        return self.local_energy()

    def roofline_items(self):
        """Algorithmic HBM bytes per step: A^-1 (B nfp^2 FP64) read and, for
        the sweep, written once per sweep; configs and the per-walker rows."""
        B, N, nf = self.B, self.N, self.nf
        nfp = self.k.slater_nfp(nf)
        sz = 4 if self.fp32 else 8
        return {
            "sweep": ("hbm", self.spp * B * (2 * nfp * nfp * sz + 2 * 2 * N + 2 * 8 * N)),
            "local_energy": ("hbm", B * (nfp * nfp * sz + 2 * N + 8 * N + 8)),
            "logderiv": ("hbm", B * (nf * nfp * sz + 2 * N + self.Np * 8)),
            "refresh": ("hbm", B * (nfp * nfp * sz + 2 * N + 8 * N)),
        }

    def oracle_params(self):
        return [self.M.cpu().numpy()], 0.0


class NnbfRun:
    """NNBF walkers through the reference-shaped model class NNBF_GPU."""

    names = ["sweep", "local_energy", "logderiv", "sr", "update"]
    factored = False

    def __init__(self, w, a, dev, H, fxs):
        import torch

        from vmc_torch_amd import NNBF_GPU, kernels
        self.k, self.a, self.w, self.H, self.fxs = kernels, a, w, H, fxs
        N, nf, B = w["Lx"] * w["Ly"], w["nu"] + w["nd"], w["B"]
        self.N, self.nf, self.B = N, nf, B
        self.mlp = a.mlp or w["mlp"]
        self.det = getattr(a, "det", None) or "fp64"
        self.model = NNBF_GPU(N, nf, hidden_dim=w["hidden"], activation="relu",
                              bf_scale=w["bf_scale"], gemm=self.mlp, init="xavier", m_std=0.1,
                              generator=torch.Generator().manual_seed(0),
                              chunk=a.chunk, det=self.det).to(dev)
        self.Np = self.model.n_params
        self.e_sweep = H.graph.edge_tensor(dev, "sweep")
        self.e_graph = H.graph.edge_tensor(dev, "graph")
        self.proposals_per_step = self.e_sweep.shape[0]
        self.lp = None
        if a.sampler == "learned":
            from vmc_torch_amd.sampler import LearnedProposalSamplerGPU
            self.lp = LearnedProposalSamplerGPU(w["Lx"], w["Ly"], channels=32, seed=0x1EA,
                                                generator=torch.Generator().manual_seed(0),
                                                device=dev, moves_per_proposal=a.lp_moves)
            self.proposals_per_step = (a.proposals_per_sample if a.proposals_per_sample > 0
                                       else max(1, self.e_sweep.shape[0] // 4))
            self.n_steps = 0
        self.o_dtype = a.o_dtype or w.get("o_dtype", "fp32")
        o_dtype = torch.float64 if self.o_dtype == "fp64" else torch.float32
        self.factored = (a.gram or w["gram"]) == "factored"
        self.dp = torch.empty(self.Np, dtype=torch.float64, device=dev)
        if self.factored:
            H = w["hidden"]
            self.factor_widths = [2 * N * nf, 2 * N, H, H, H, H]
            self.F = torch.empty(B, sum(self.factor_widths), dtype=torch.float64, device=dev)
            self.O = self.F
            self.ws = self.k.SRFactoredWorkspace()
        else:
            self.O = torch.empty(B, self.Np, dtype=o_dtype, device=dev)
            self.ws = self.k.SRWorkspace()

    def burn(self, n, seed, off):
        self.model.mcmc_sweep(self.fxs, self.e_sweep, n, seed, off)
        if self.lp is not None:       # train the proposal net on the chain's walkers
            for _ in range(100):
                self.lp.sweeps(self.fxs, self.model, 2)
                self.lp.train_step(self.fxs[:1024], self.model, lr=0.05)

    def local_energy(self):
        return self.model.local_energy(self.fxs, self.H)

    def _sample(self, seed, off):
        """One sample per walker; returns the accepted moves per walker."""
        if self.lp is None:
            acc0 = self.model.walker_state(self.fxs)["n_accept"].clone()
            _, _, nacc = self.model.mcmc_sweep(self.fxs, self.e_sweep, 1, seed, off)
            return nacc - acc0
        self.lp.sweeps(self.fxs, self.model, self.proposals_per_step)
        return self.lp.last_acceptance[0]

    def step(self, mark, seed, off, gram, phases):
        m = [mark()]
        nacc = self._sample(seed, off)
        m.append(mark())
        E = self.local_energy()
        m.append(mark())
        if self.factored:
            FO, _ = self.model.log_derivative_factored(self.fxs, out=self.F)
            m.append(mark())
            _, stats, info = self.k.sr_minsr_factored(FO, E, None, self.w["shift"], self.ws,
                                                      dp=self.dp, phase_ms=phases)
        else:
            self.model.log_derivative(self.fxs, dtype=self.O.dtype, out=self.O)
            m.append(mark())
            _, stats, info = self.k.sr_minsr(self.O, E, None, self.w["shift"], gram,
                                             0 if gram == "fp64" else 20, 1e-8, self.ws,
                                             dp=self.dp, phase_ms=phases, inplace=True)
        m.append(mark())
        off_p = 0
        for p in self.model.params:
            n = p.numel()
            self.k.sgd_update(p.data.view(-1), self.dp[off_p:off_p + n], self.w["lr"])
            off_p += n
        self.model.invalidate()
        if self.lp is not None:
            self.n_steps += 1
            if self.n_steps % 10 == 0:
                self.lp.train_step(self.fxs[:1024], self.model, lr=0.05)
        m.append(mark())
        return m, E, nacc, stats, info

    def pre_tau(self):
        pass

    def tau_sample(self, seed, off):
        self._sample(seed, off)
        return self.local_energy()

    def roofline_items(self):
        """Sweep and local energy: dense-layer flops (layer 2 and the output
        layer; the first layer is a gather) per amplitude evaluated, one per
        proposal and one per connected configuration.  logderiv: the O write."""
        H, nD = self.w["hidden"], 2 * self.N * self.nf
        mlp = 2.0 * (H * H + H * nD)
        n_prop = self.proposals_per_step * self.B
        x = self.fxs
        e = self.e_graph.long()
        n_conn = 0
        for bit in (1, 2):                       # down, up occupation bits
            occ = (x & bit) != 0
            n_conn += int((occ[:, e[:, 0]] != occ[:, e[:, 1]]).sum().item())
        self.n_conn = n_conn
        return {
            "sweep": ("mfma_bf16", n_prop * mlp),
            "local_energy": ("mfma_bf16", n_conn * mlp),
            "logderiv": ("hbm", float(self.O.numel() * self.O.element_size())),
        }

    def oracle_params(self):
        # the FP64 master parameters: the oracle is FP64 throughout, so the
        # difference includes the rounding of the FP32 / BF16 modes
        return [p.detach().cpu().numpy() for p in self.model.params], self.w["bf_scale"]


class PepsRun(NnbfRun):
    """fPEPS / NN-fTNS walkers through the reference-shaped model classes."""

    def __init__(self, w, a, dev, H, fxs):
        import torch

        from vmc_torch_amd import NNfTNS_Model_GPU, fPEPS_Model_GPU, fPEPS_Model_reuse_GPU, kernels
        self.k, self.a, self.w, self.H, self.fxs = kernels, a, w, H, fxs
        N, nf, B = w["Lx"] * w["Ly"], w["nu"] + w["nd"], w["B"]
        self.N, self.nf, self.B = N, nf, B
        g = torch.Generator().manual_seed(0)
        chi = w.get("chi", -1)
        if w["hidden"] > 0:
            self.model = NNfTNS_Model_GPU(w["Lx"], w["Ly"], w["D"], hidden_dim=w["hidden"],
                                          nn_eta=w["eta"], generator=g, amp=a.peps_amp,
                                          max_bond=chi).to(dev)
        else:
            cls = fPEPS_Model_reuse_GPU if getattr(a, "peps_reuse", False) else fPEPS_Model_GPU
            kw = {"amp": a.peps_amp} if a.peps_amp != "fp64" else {}
            self.model = cls(w["Lx"], w["Ly"], w["D"], generator=g, max_bond=chi,
                             **kw).to(dev)
        self.Np = self.model.n_params
        self.e_sweep = H.graph.edge_tensor(dev, "sweep")
        self.e_graph = H.graph.edge_tensor(dev, "graph")
        gram = a.gram or w["gram"]
        self.factored = gram == "factored"
        self.dp = torch.empty(self.Np, dtype=torch.float64, device=dev)
        if self.factored:
            P, Hd = self.model.n_tn_params, w["hidden"]
            self.factor_widths = [P, 2 * N, Hd, Hd, Hd, Hd] if Hd > 0 else [P]
            self.F = torch.empty(B, sum(self.factor_widths), dtype=torch.float64, device=dev)
            self.O = self.F
            self.ws = self.k.SRFactoredWorkspace()
        else:
            odt = torch.float64 if gram == "fp64" else torch.float32
            self.O = torch.empty(B, self.Np, dtype=odt, device=dev)
            self.ws = self.k.SRWorkspace()
        self.lp, self.proposals_per_step = None, self.e_sweep.shape[0]

    def roofline_items(self):
        """Sweep and local energy: FP64 flops of every amplitude evaluated
        (column-pair contraction, or the boundary MPS's dense products without
        its eigensolver; the MLP when present).  logderiv: the O write."""
        D, Ly, Hd = self.w["D"], self.w["Ly"], self.w["hidden"]
        if getattr(self.model, "_bmps", False):
            flops = bmps_amplitude_flops(self.w["Lx"], Ly, D, self.w["chi"])
        else:
            contr = 0
            for y in range(Ly):
                dl = 1 if y == 0 else D
                dr = 1 if y == Ly - 1 else D
                da, da2 = dl * dl, dr * dr
                contr += 2 * (da * D * da2 * D) + da2 * D * da2 * da + da2 * da2 * da2 * D
            flops = 2.0 * contr
        if Hd:
            flops += 2.0 * (2 * self.N * Hd + Hd * Hd + Hd * self.model.n_tn_params)
        n_prop = self.e_sweep.shape[0] * self.B
        x, e = self.fxs, self.e_graph.long()
        n_conn = 0
        for bit in (1, 2):
            occ = (x & bit) != 0
            n_conn += int((occ[:, e[:, 0]] != occ[:, e[:, 1]]).sum().item())
        return {
            "sweep": ("fp64", n_prop * flops),
            "local_energy": ("fp64", n_conn * flops),
            "logderiv": ("hbm", self.B * self.Np * float(self.O.element_size())),
        }

    def oracle_params(self):
        ps = [p.detach().cpu().numpy() for p in self.model.params]
        chi = self.w.get("chi", 0) if getattr(self.model, "_bmps", False) else 0
        return ps, {"D": self.w["D"], "eta": self.w["eta"], "chi": chi}


def bmps_amplitude_flops(Lx, Ly, D, chi):
    """FP64 flops of the dense products of one boundary-MPS amplitude (the
    right Grams, C, C R, rho = C R C^T and the projection per site of every
    truncated row, the absorbed tensors of the others, the join), bonds
    taken at their bound min(chi, rank); the Jacobi eigensolver is not
    counted."""
    h = Lx // 2
    hd = [1] + [D] * (Ly - 1) + [1]
    total = 0
    for nrows in (h, Lx - h):
        bd = list(hd)
        for _ in range(nrows - 1):
            nbd = [b * x for b, x in zip(bd, hd)]
            if max(nbd) <= chi:
                total += sum(nbd[y] * D * nbd[y + 1] * D for y in range(Ly))
                bd = nbd
                continue
            for y in range(1, Ly):      # Grams: A_v R A_v^T for every v
                n, m = nbd[y], nbd[y + 1]
                total += D * (n * m * D + n * m * m + n * n * m)
            k = 1
            new = [1]
            for y in range(Ly):
                m = nbd[y + 1]
                total += k * hd[y] * D * bd[y + 1] * bd[y] + k * D * m * hd[y] * D
                if y < Ly - 1:
                    nC = k * D
                    kn = min(chi, nC)
                    if nC > chi:
                        total += nC * m * m + nC * nC * m // 2 + kn * nC * m
                    k = kn
                new.append(k)
            bd = new[:Ly] + [1]
    total += Ly * chi * chi * D * chi * 2         # join
    return 2.0 * total


# --------------------------------------------------------------------- main
def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU "
                                   "sharding is designed but not part of this build",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if WORKLOADS[a.workload]["kind"] == "mixed_lp":
        return main_mixed_lp(a, WORKLOADS[a.workload])
    if a.sampler == "learned" and WORKLOADS[a.workload]["kind"] != "nnbf":
        print(json.dumps({"error": "--sampler learned is wired for the NNBF workloads"}))
        sys.exit(2)
    import numpy as np
    import torch

    from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch
    from vmc_torch_amd.stats import tau_int

    w = WORKLOADS[a.workload]
    gram = a.gram or w["gram"]
    if gram == "factored" and w["kind"] not in ("nnbf", "peps"):
        print(json.dumps({"error": "--gram factored is wired for the NNBF and fPEPS workloads"}))
        sys.exit(2)
    Lx, Ly, nu, nd, B = w["Lx"], w["Ly"], w["nu"], w["nd"], w["B"]
    N, nf = Lx * Ly, nu + nd
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.copy_probe_mib > 0:
        src = torch.ones(a.copy_probe_mib << 17, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {a.copy_probe_mib} MiB device-to-device")
        del src, dst
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nf,
                                                   n_fermions_per_spin=(nu, nd))
    H.precompute_hops_gpu(dev)
    # walkers: the reference's initial pattern, permuted per walker (seeded)
    base = torch.tensor([1, 2] * (N // 2), dtype=torch.int64)
    base[: N - nf] = 0
    perms = torch.argsort(torch.rand(B, N, generator=torch.Generator().manual_seed(42)), dim=1)
    fxs = base[perms].contiguous().to(dev)
    fxs0 = fxs.clone()
    run = {"slater": SlaterRun, "nnbf": NnbfRun, "peps": PepsRun}[w["kind"]](w, a, dev, H, fxs)
    seed, sweep_no = 0x5EED, 0
    spp = getattr(run, "spp", 1)
    if a.sweeps_per_sample != 1 and w["kind"] != "slater":
        print(json.dumps({"error": "--sweeps-per-sample is wired for the Slater workloads"}))
        sys.exit(2)
    log(f"{a.workload}: {B} walkers, {run.Np} parameters; burn-in {w['burn']} sweeps")
    run.burn(w["burn"], seed, 0)
    sweep_no += w["burn"]
    torch.cuda.synchronize()
    log("burn-in done")

    def mark():
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    names = run.names
    ev = {k: [] for k in names}
    E_hist, accs = [], []
    for _ in range(a.warmup):
        run.step(mark, seed, sweep_no, gram, None)
        sweep_no += spp
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        m, E, acc, stats, info = run.step(mark, seed, sweep_no, gram, None)
        sweep_no += spp
        for k, x0, x1 in zip(names, m[:-1], m[1:]):
            ev[k].append((x0, x1))
        E_hist.append(E.clone())
        accs.append(acc)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / a.steps
    log(f"{a.steps} timed SR iterations: {ms_step:.2f} ms each")
    info_v = int(info.item())
    st_v = stats.cpu().tolist()
    kt = {k: sum(x0.elapsed_time(x1) for x0, x1 in ev[k]) / a.steps for k in names}  # ms
    # SR phase split, measured on extra (untimed) steps: the phase timer
    # synchronises the host once per SR call
    sr_phases = []
    for _ in range(3):
        ph = []
        run.step(mark, seed, sweep_no, gram, ph)
        sweep_no += spp
        sr_phases.append(ph)
    torch.cuda.synchronize()
    phase_names = ["center", "gram", "factor", "solve", "dp_stats", "chol_diag_launch"]
    if gram == "factored":     # vmcx_sr_minsr_factored reports the Grams first
        phase_names = ["gram", "center", "factor", "solve", "dp_stats", "chol_diag_launch"]
    phase_ms = {k: sum(p[i] for p in sr_phases) / len(sr_phases)
                for i, k in enumerate(phase_names)}
    acc_rate = (float(torch.stack(accs).double().mean())
                / getattr(run, "proposals_per_step", run.e_sweep.shape[0]))
    e_site = [float(v) / N for v in torch.stack(E_hist).mean(dim=1)]

    # tau_int of E_loc at fixed parameters, one sample per sweep (as in the loop)
    run.pre_tau()
    series = []
    for _ in range(a.tau_sweeps):
        series.append(run.tau_sample(seed, sweep_no).clone())
        sweep_no += spp
    torch.cuda.synchronize()
    tau = tau_int(torch.stack(series, dim=1))
    log(f"SR phases and tau_int ({a.tau_sweeps} sweeps) done: tau_int = {tau:.3f}")
    sr_iters = a.steps / wall
    samples_per_s = B * sr_iters
    ess_per_s = samples_per_s / tau

    # roofline of every component: algorithmic work per step / mean duration
    Np, Ns = run.Np, B
    kern = {}
    for k, (bound, work) in run.roofline_items().items():
        secs = kt[k] * 1e-3
        if bound == "hbm":
            kern[k] = {"ms": kt[k], "bound": "hbm", "achieved": work / secs / 1e9,
                       "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": work / secs / PEAK_HBM}
        elif bound == "fp64":
            kern[k] = {"ms": kt[k], "bound": "mfma", "achieved": work / secs / 1e12,
                       "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                       "frac": work / secs / PEAK_FP64,
                       "note": "FP64 flops of every amplitude evaluated (contraction + MLP)"}
        else:
            kern[k] = {"ms": kt[k], "bound": "mfma", "achieved": work / secs / 1e12,
                       "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                       "frac": work / secs / PEAK_BF16,
                       "note": "dense-layer flops of every amplitude evaluated, BF16 operands"}
    peak_sr = {"fp64": PEAK_FP64, "fp32": PEAK_FP32_MFMA, "bf16": PEAK_BF16,
               "bf16x3": PEAK_BF16, "factored": PEAK_FP64}[gram]
    if gram == "factored":
        # one symmetric product per factor: Ns^2 k_f, k_f = 2N nf, 2N, 4 x hidden
        gram_flops = float(Ns) * Ns * sum(run.factor_widths)
    else:
        gram_flops = float(Ns) * Ns * Np * {"fp64": 1, "fp32": 2, "bf16": 2, "bf16x3": 4}[gram]
    chol_flops = float(Ns) ** 3 / 3
    kern["sr"] = {"ms": kt["sr"], "phases_ms": phase_ms}
    kern["update"] = {"ms": kt["update"]}

    def flop_roof(flops, ms, peak, note):
        return {"ms": ms, "bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12,
                "peak": peak / 1e12, "unit": "TFLOP/s", "frac": flops / (ms * 1e-3) / peak,
                "note": note}

    def hbm_roof(nbytes, ms):
        return {"ms": ms, "bound": "hbm", "achieved": nbytes / (ms * 1e-3) / 1e9,
                "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": nbytes / (ms * 1e-3) / PEAK_HBM}
    kern["sr.gram"] = flop_roof(gram_flops, phase_ms["gram"], peak_sr,
                                "algorithmic flops Ns^2 Np (FP64, symmetric product), "
                                "2 Ns^2 Np (reduced-precision GEMM) or 4 Ns^2 Np (bf16x3: "
                                "two full products), Ns^2 sum(k_f) (factored: one "
                                "symmetric product per factor of O); FP64: the matrix-core "
                                "SYRK kernel (one workgroup per 128x128 lower tile, the "
                                "reduction split when the triangle has few tiles) plus the "
                                "pass that adds the split partials")
    kern["sr.factor"] = flop_roof(chol_flops, phase_ms["factor"], PEAK_FP64,
                                  "Ns^3/3: blocked FP64 Cholesky (LDS diagonal kernel, "
                                  "panel GEMM, matrix-core SYRK per two panels)")
    osz = run.O.element_size()
    if gram == "factored":
        # combination: 7 factor-Gram triangles read, T written; row sums, the
        # centring pass (read + write) and ||g|| read T; T copied once (full)
        tri = Ns * (Ns + 1) / 2 * 8.0
        n_reads = 7 if len(run.factor_widths) > 1 else 1
        cbytes = (n_reads + 1 + 1 + 2 + 1) * tri + 2.0 * Ns * Ns * 8
    elif gram == "fp64":
        cbytes = 2.0 * Ns * Np * osz + Ns * Np * osz          # read O twice, write X
    else:
        cbytes = 2.0 * Ns * Np * osz + Ns * Np * (osz + (4 if gram == "bf16x3" else 2))
    kern["sr.center"] = hbm_roof(cbytes, phase_ms["center"])
    # one launch of the diagonal-block Cholesky kernel: Nb^3/3 (factor) +
    # Nb^3/3 (block inverse) flops, Nb = 128; Ns/128 launches per SR call
    nb = 128
    n_diag = (Ns + nb - 1) // nb
    dl_ms = phase_ms["chol_diag_launch"]
    kern["sr.chol_diag"] = dict(
        flop_roof(2.0 * nb ** 3 / 3, dl_ms, PEAK_FP64,
                  "one launch: 128x128 FP64 Cholesky + triangular inverse in LDS "
                  "(latency bound)"), launches_per_step=n_diag, per_step_ms=dl_ms * n_diag)
    if w["kind"] == "slater" and gram == "fp64":
        kern["sr.factor"]["library_same_T"] = library_factor_times(run.O, w["shift"])
    # the roofline line is the kernel (or single library call) with the
    # largest time per SR iteration
    single = {"sr.gram": 1, "sr.chol_diag": n_diag}
    if w["kind"] == "slater":
        single.update({"sweep": 1, "local_energy": 1, "logderiv": 1, "refresh": 1})
    elif run.factored or w["kind"] == "peps":
        # with the Gram formed from the factors (or the small fPEPS O) the SR
        # step no longer sets the step: the sampling components (several
        # kernels each) compete
        single.update({"sweep": 1, "local_energy": 1, "logderiv": 1})
    for k, n in single.items():
        kern[k].setdefault("launches_per_step", n)
        kern[k].setdefault("per_step_ms", kern[k]["ms"])
    dom = max(single, key=lambda k: kern[k]["per_step_ms"])
    roof = {kk: kern[dom][kk] for kk in ("bound", "achieved", "peak", "unit", "frac")}
    roof["kernel"] = dom
    # HBM bytes per call of the roofline kernel from rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs scaled by a 1 GiB copy (profiles/pmc_traffic_*.md);
    # null where no such measurement exists for the workload and kernel
    roof["traffic"] = TRAFFIC.get((a.workload, dom, gram))

    if w["kind"] == "slater":
        dtype = "f64" if gram == "fp64" else f"f64+{gram}-gram"
    elif w["kind"] == "peps":
        dtype = "f64" if gram == "fp64" else f"f64/f32-O/{gram}-gram/f64-cholesky"
        if gram == "factored":
            dtype = "f64/factors-of-O/f64-gram/f64-cholesky"
    else:
        o = "f64" if run.o_dtype == "fp64" else "f32"
        dtype = f"{run.mlp}-gemm/f64-det/{o}-O/{gram}-gram/f64-cholesky"
        if gram == "factored":
            dtype = f"{run.mlp}-gemm/f64-det/f64-factors-of-O/f64-gram/f64-cholesky"
        if a.sampler == "learned":
            dtype += "/bf16-proposal-net"
    out = {
        "metric": "effective samples/s + SR iters/s at 8x8 Hubbard; |dE/site| vs FP64 oracle",
        "value": ess_per_s,
        "unit": "effective samples/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded parameters, seeded initial walkers)",
        "config": {"workload": a.workload, "description": w["desc"], "walkers": B,
                   "samples_per_sr_iter": Ns, "n_params": Np, "burn_in_sweeps": w["burn"],
                   "gram": gram, "diag_shift": w["shift"], "lr": w["lr"],
                   "sampler": a.sampler, "sweeps_per_sample": spp,
                   **({"moves_per_proposal": a.lp_moves} if a.sampler == "learned" else {}),
                   "proposals_per_sample": getattr(run, "proposals_per_step", None),
                   "parallelism": "walkers on 1 GPU",
                   **({"environment_reuse": bool(a.peps_reuse)}
                      if w["kind"] == "peps" and w.get("hidden", 0) == 0 else {}),
                   **({"amplitudes": a.peps_amp} if w["kind"] == "peps" else {})},
        "sr_iters_per_s": sr_iters,
        "walker_sweeps_per_s": samples_per_s * spp,
        "sweep_only_walker_sweeps_per_s": B * spp / (kt["sweep"] * 1e-3),
        "tau_int": tau, "acceptance": acc_rate,
        "energy_per_site_first_last": [e_site[0], e_site[-1]],
        "sr_solver": {"info": info_v, "rel_residual": st_v[2], "grad_norm": st_v[3],
                      "pcg_iters": st_v[4]},
        "roofline": roof,
        "kernels": kern,
    }
    if w["kind"] == "slater":
        out["config"]["refresh_every"] = a.refresh_every
        out["config"]["precision"] = a.slater_precision
        if a.slater_precision == "fp32":
            out["dtype"] = "f32-state/f64-refresh/f64-sr"
    if a.workload == "hubbard8x8_slater" and a.slater_precision == "fp64":
        # the FP32 form of the same configuration (the FP64 line is the headline)
        out["fp32_form"] = slater_fp32_form(a, w, fxs0, dev, H, a.steps, a.warmup)
        log(f"FP32 form: {out['fp32_form']['ms_per_step']:.2f} ms per SR iteration")
        if not a.no_target:
            out["target"] = nnbf_target(a, dev, a.target_steps)

    if not a.no_cpu_baseline and spp == 1:   # the CPU sample is one sweep per sample
        # FP64 oracle on a subset of the final walkers: |dE/site|, then the timed baseline
        from oracle import fpeps as ofp
        from oracle import hubbard as ohub
        from oracle import nnbf as onn
        from oracle import slater as osl
        params, bf = run.oracle_params()
        n_chk = 128 if w["kind"] == "slater" else 32
        xs = fxs[:n_chk].cpu().numpy()
        Eg = run.local_energy()[:n_chk].cpu().numpy()
        Ho = ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd)
        if w["kind"] == "slater":
            Eo = osl.local_energy(Ho, params[0], xs)
        elif w["kind"] == "nnbf":
            Eo = onn.local_energy(Ho, params, xs, 2, bf)
        else:
            Eo = ofp.local_energy(Ho, ofp.PEPSShape(Lx, Ly, bf["D"]), params[0], xs,
                                  params[1:] or None, bf["eta"], chi=bf.get("chi", 0))
        out["dE_per_site_vs_fp64_oracle"] = {
            "mean": float(abs(Eg.mean() - Eo.mean()) / N),
            "max_per_sample": float(np.abs(Eg - Eo).max() / N), "samples": int(len(xs))}
        log("FP64 oracle check done; CPU baseline starts")
        out["cpu_baseline"] = cpu_baseline(w["kind"], Lx, Ly, nu, nd, params, bf,
                                           fxs[:4096].cpu().numpy(), Ns, tau, w["shift"],
                                           a.cpu_seconds)
    print(json.dumps(out))



# ------------------------------------------------ forms beside the headline
def measure_form(run, w, a, gram, steps, warmup, seed, sweep_no, tau_sweeps, tau=None):
    """Time `steps` SR iterations of a prepared run (after `warmup` untimed
    ones), then tau_int of E_loc over `tau_sweeps` samples at fixed
    parameters (skipped when tau is given: a form with the same chain)."""
    import torch

    from vmc_torch_amd.stats import tau_int

    def mark():
        return None

    def nomark():
        return None
    for _ in range(warmup):
        run.step(nomark, seed, sweep_no, gram, None)
        sweep_no += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    accs = []
    for _ in range(steps):
        _, E, acc, stats, info = run.step(mark, seed, sweep_no, gram, None)
        sweep_no += 1
        accs.append(acc)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if tau is None:
        run.pre_tau()
        series = []
        for _ in range(tau_sweeps):
            series.append(run.tau_sample(seed, sweep_no).clone())
            sweep_no += 1
        torch.cuda.synchronize()
        tau = tau_int(torch.stack(series, dim=1))
    B = run.B
    sr = steps / wall
    return {"ms_per_step": 1e3 * wall / steps, "sr_iters_per_s": sr, "tau_int": tau,
            "effective_samples_per_s": B * sr / tau,
            "acceptance": float(torch.stack(accs).double().mean()) / run.proposals_per_step,
            "sr_info": int(info.item()), "energy_per_site": float(E.mean()) / run.N}, sweep_no


def slater_fp32_form(a, w, fxs0, dev, H, steps, warmup):
    """The FP32 form of the headline configuration on the same initial
    walkers: FP32 parameters and A^-1, FP32 ratios and Sherman-Morrison
    updates, FP64 refresh, local-energy sums and SR."""
    import argparse as ap
    import copy

    import numpy as np

    from oracle import hubbard as ohub
    from oracle import slater as osl
    a32 = ap.Namespace(**vars(copy.deepcopy(a)))
    a32.slater_precision = "fp32"
    run = SlaterRun(w, a32, dev, H, fxs0.clone())
    seed = 0x5EED
    run.burn(w["burn"], seed, 0)
    res, sweep_no = measure_form(run, w, a32, "fp64", steps, warmup, seed, w["burn"],
                                 a.tau_sweeps)
    # the sweep alone, both forms, on the walkers of the FP32 run (HIP events)
    import torch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    run._sweep(4, seed, sweep_no)
    ev[1].record()
    torch.cuda.synchronize()
    res["sweep_ms"] = ev[0].elapsed_time(ev[1]) / 4
    xs = run.fxs[:128].cpu().numpy()
    Eg = run.local_energy()[:128].cpu().numpy()
    Eo = osl.local_energy(ohub.Hubbard(w["Lx"], w["Ly"], 1.0, 8.0, w["nu"], w["nd"]),
                          run.M.cpu().numpy(), xs)
    N = w["Lx"] * w["Ly"]
    res["dE_per_site_vs_fp64_oracle"] = {"mean": float(abs(Eg.mean() - Eo.mean()) / N),
                                         "max_per_sample": float(np.abs(Eg - Eo).max() / N),
                                         "samples": int(len(xs))}
    res["dtype"] = "f32-state/f64-refresh/f64-sr"
    return res


def nnbf_target(a, dev, steps):
    """The mixed-precision target comparison on 8x8 U=8 with the NNBF ansatz (128 -> 256
    -> 256 -> 8192, 2,212,352 parameters, 4096 walkers): the plain FP64 GPU
    form (FP64 network and determinant, plain sweep sampler) with the exact
    FP64 Gram formed from the factors of O, the same with O formed (FP64 Gram
    of the 4096 x 2.2 M matrix, for reference), and the mixed-precision form
    (BF16 network with FP32 accumulation on the matrix cores, FP32
    determinants of the amplitudes from entries formed in FP64, learned
    proposal, the same factored FP64 SR; O_k from FP64 inverses).  One
    burn-in (plain sampler, BF16 network) serves all forms; the parameters
    start equal (one seed)."""
    import argparse as ap
    import copy

    import numpy as np
    import torch

    from oracle import hubbard as ohub
    from oracle import nnbf as onn
    from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch
    w = WORKLOADS["hubbard8x8_nnbf"]
    Lx, Ly, nu, nd, B = w["Lx"], w["Ly"], w["nu"], w["nd"], w["B"]
    N, nf = Lx * Ly, nu + nd
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nf,
                                                   n_fermions_per_spin=(nu, nd))
    H.precompute_hops_gpu(dev)
    base = torch.tensor([1, 2] * (N // 2), dtype=torch.int64)
    base[: N - nf] = 0
    perms = torch.argsort(torch.rand(B, N, generator=torch.Generator().manual_seed(42)), dim=1)
    fxs0 = base[perms].contiguous().to(dev)

    def args(**kw):
        x = ap.Namespace(**vars(copy.deepcopy(a)))
        x.proposals_per_sample, x.chunk = 0, 65536
        for k, v in kw.items():
            setattr(x, k, v)
        return x
    seed = 0x5EED
    burn_run = NnbfRun(w, args(mlp="bf16", gram="factored", o_dtype=None, sampler="plain",
                               det="fp64"), dev,
                       H, fxs0)
    burn_run.burn(w["burn"], seed, 0)
    torch.cuda.synchronize()
    log(f"target: 8x8 NNBF burn-in ({w['burn']} sweeps) done")
    forms = {}
    specs = [("fp64_factored", dict(mlp="fp64", gram="factored", o_dtype=None,
                                    sampler="plain", det="fp64"), steps, None),
             ("fp64_O_formed", dict(mlp="fp64", gram="fp64", o_dtype="fp64", sampler="plain",
                                    det="fp64"),
              max(2, steps // 3), "fp64_factored"),
             ("bf16_learned", dict(mlp="bf16", gram="factored", o_dtype=None,
                                   sampler="learned", det="fp32"), steps, None)]
    bf_run = None
    for name, kw, n_steps, same_chain in specs:
        ai = args(**kw)
        run = NnbfRun(w, ai, dev, H, fxs0.clone())
        if run.lp is not None:
            run.burn(0, seed, 0)              # trains the proposal net (100 iterations)
        gram = ai.gram
        tau = forms[same_chain]["tau_int"] if same_chain else None
        res, _ = measure_form(run, w, ai, gram, n_steps, 1, seed, w["burn"], a.tau_sweeps, tau)
        res["mlp"], res["gram"], res["sampler"] = ai.mlp, gram, ai.sampler
        res["det"] = getattr(ai, "det", None) or "fp64"
        if same_chain:
            res["tau_int_note"] = f"the chain of {same_chain} (same sampler and arithmetic)"
        forms[name] = res
        log(f"target form {name}: {res['ms_per_step']:.1f} ms per SR iteration, "
            f"tau_int {res['tau_int']:.2f}")
        if name == "bf16_learned":
            bf_run = run
        else:
            del run
        torch.cuda.empty_cache()
    base_f, mixed = forms["fp64_factored"], forms["bf16_learned"]
    out = {"workload": "hubbard8x8_nnbf", "description": w["desc"].replace(
               "BF16 GEMMs", "network precision per form"),
           "walkers": B, "forms": forms,
           "ess_ratio_vs_fp64_factored": mixed["effective_samples_per_s"]
           / base_f["effective_samples_per_s"],
           "sr_ratio_vs_fp64_factored": mixed["sr_iters_per_s"] / base_f["sr_iters_per_s"],
           "ess_ratio_vs_fp64_O_formed": mixed["effective_samples_per_s"]
           / forms["fp64_O_formed"]["effective_samples_per_s"],
           "sr_ratio_vs_fp64_O_formed": mixed["sr_iters_per_s"]
           / forms["fp64_O_formed"]["sr_iters_per_s"],
           "target": {"ess_ratio": 3.0, "sr_ratio": 2.0, "dE_per_site": 1e-4}}
    # |dE/site| of the mixed form against the FP64 oracle, 1024 of its walkers
    n_chk = 1024
    xs = bf_run.fxs[:n_chk].cpu().numpy()
    Eg = bf_run.local_energy()[:n_chk].cpu().numpy()
    params, bf = bf_run.oracle_params()
    Eo = onn.local_energy(ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd), params, xs, 2, bf, chunk=64)
    d = np.abs(Eg - Eo) / N
    out["dE_per_site_vs_fp64_oracle"] = {"estimate": float(abs(Eg.mean() - Eo.mean()) / N),
                                         "per_sample_mean": float(d.mean()),
                                         "per_sample_max": float(d.max()),
                                         "samples": int(n_chk)}
    log("target: FP64 oracle check of the mixed form done")
    return out


# ------------------------------------------------- batched learned proposal
def lp_logits_flops(N, E, C):
    """FP32 flops of one net evaluation: three 3x3 convolutions (2 -> C, C -> C,
    C -> C; every tap counted, including the zero-padding ones) and the head
    (3 move types x 2C per edge)."""
    return 2.0 * N * 9 * (2 * C + 2 * C * C) + 2.0 * 3 * E * 2 * C


def main_mixed_lp(a, w):
    """The learned proposal on an uneven batch of lattices against the plain
    sweep sampler, on the same wavefunctions and walkers."""
    import numpy as np
    import torch

    from vmc_torch_amd import NNBF_GPU, kernels
    from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch
    from vmc_torch_amd.sampler import (LearnedProposalSamplerGPU,
                                       MetropolisExchangeSpinfulSamplerGPU)
    from vmc_torch_amd.stats import tau_int
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    C = w["channels"]
    net = LearnedProposalSamplerGPU.init_params(C, torch.Generator().manual_seed(0)).to(dev)
    inst = []
    for n_i, (Lx, Ly, nu, nd, B) in enumerate(w["instances"]):
        N, nf = Lx * Ly, nu + nd
        H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nf,
                                                       n_fermions_per_spin=(nu, nd))
        H.precompute_hops_gpu(dev)
        model = NNBF_GPU(N, nf, hidden_dim=w["hidden"], activation="relu",
                         bf_scale=w["bf_scale"], gemm=w["mlp"], init="xavier", m_std=0.1,
                         generator=torch.Generator().manual_seed(n_i)).to(dev)
        base = torch.tensor([1, 2] * (N // 2), dtype=torch.int64)
        base[: N - nf] = 0
        perms = torch.argsort(torch.rand(B, N, generator=torch.Generator().manual_seed(42 + n_i)),
                              dim=1)
        E = H.graph.edge_tensor(dev, "sweep").shape[0]
        inst.append(dict(
            name=f"{Lx}x{Ly}", Lx=Lx, Ly=Ly, nu=nu, nd=nd, B=B, N=N, nf=nf, H=H, model=model,
            x=base[perms].contiguous().to(dev), E=E, k=max(1, E // 4),
            plain=MetropolisExchangeSpinfulSamplerGPU(seed=0x5EED + n_i),
            lp=LearnedProposalSamplerGPU(Lx, Ly, channels=C, seed=0x1EA + n_i, params=net)))
    Btot = sum(I["B"] for I in inst)
    log(f"{a.workload}: {Btot} walkers on {len(inst)} lattices; burn-in {w['burn']} sweeps")
    for I in inst:
        I["plain"].sweeps(I["x"], I["model"], I["H"].graph, w["burn"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = []
    for it in range(w["train_iters"]):
        for I in inst:
            I["lp"].sweeps(I["x"], I["model"], 2)
            loss = I["lp"].train_step(I["x"][:w["train_walkers"]], I["model"], lr=w["train_lr"])
            if it in (0, w["train_iters"] - 1):
                losses.append(loss)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    log(f"burn-in and {w['train_iters']} training iterations done ({t_train:.1f} s)")

    def plain_step(I):
        I["plain"].sweeps(I["x"], I["model"], I["H"].graph, 1)
        return I["model"].local_energy(I["x"], I["H"])

    def lp_step(I):
        I["lp"].sweeps(I["x"], I["model"], I["k"])
        return I["model"].local_energy(I["x"], I["H"])

    def form(name):
        step = plain_step if name == "plain" else lp_step
        for _ in range(a.warmup):
            for I in inst:
                step(I)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s_ in range(a.steps):
            for I in inst:
                step(I)
            if name == "learned" and (s_ + 1) % w["train_every"] == 0:
                for I in inst:
                    I["lp"].train_step(I["x"][:w["train_walkers"]], I["model"],
                                       lr=w["train_lr"])
        torch.cuda.synchronize()
        t_step = (time.perf_counter() - t0) / a.steps
        # tau_int of E_loc at fixed parameters and a fixed net, one sample per step
        per = []
        for I in inst:
            smp = I["plain"] if name == "plain" else I["lp"]
            series, acc = [], 0.0
            for _ in range(a.tau_sweeps):
                series.append(step(I) / I["N"])
                n_acc, n_prop = smp.last_acceptance
                acc += float(n_acc.double().mean()) / n_prop
            ser = torch.stack(series, dim=1)
            per.append({"lattice": I["name"], "walkers": I["B"], "tau_int": tau_int(ser),
                        "acceptance_per_proposal": acc / a.tau_sweeps,
                        "proposals_per_sample": I["E"] if name == "plain" else I["k"],
                        "energy_per_site": float(ser.mean()),
                        "energy_per_site_stderr": float(ser.mean(dim=1).std()) / I["B"] ** 0.5})
        torch.cuda.synchronize()
        ess = sum(I["B"] / r["tau_int"] for I, r in zip(inst, per)) / t_step
        log(f"{name}: {1e3 * t_step:.1f} ms per step, {ess:.0f} effective samples/s, tau_int "
            + ", ".join(f"{r['lattice']} {r['tau_int']:.2f}" for r in per))
        return {"effective_samples_per_s": ess, "ms_per_step": 1e3 * t_step,
                "lattices": per}

    plain = form("plain")
    learned = form("learned")

    # per-step time of the components of the learned form, HIP events on the
    # launch stream (torch's current stream, which the library uses)
    def ev_ms(fn, reps=5):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    comp = {"lp_logits": 0.0, "amplitude": 0.0, "local_energy": 0.0}
    logits_flops, logits_launch_ms, n_launch = 0.0, 0.0, 0
    amp_flops, el_flops, amp_det, el_det = 0.0, 0.0, 0.0, 0.0
    per_lattice = {}
    for I in inst:
        t_l = ev_ms(lambda: kernels.lp_logits(net, C, I["Lx"], I["Ly"], I["x"],
                                              precision=I["lp"].precision))
        t_a = ev_ms(lambda: I["model"].forward_log(I["x"]))
        t_e = ev_ms(lambda: I["model"].local_energy(I["x"], I["H"]), 2)
        # one launch per proposal (at x'; the logits at x carry over) plus
        # the evaluation at the walkers when a sample's proposals start
        comp["lp_logits"] += (I["k"] + 1) * t_l
        comp["amplitude"] += I["k"] * t_a
        comp["local_energy"] += t_e
        per_lattice[I["name"]] = {"lp_logits_launch_ms": t_l, "amplitude_ms": t_a,
                                  "local_energy_ms": t_e}
        logits_flops += (I["k"] + 1) * I["B"] * lp_logits_flops(I["N"], I["E"], C)
        logits_launch_ms += (I["k"] + 1) * t_l
        n_launch += I["k"] + 1
        mlp = 2.0 * (w["hidden"] ** 2 + w["hidden"] * 2 * I["N"] * I["nf"])
        amp_flops += I["k"] * I["B"] * mlp
        det = 2.0 / 3.0 * I["nf"] ** 3                # FP64 LU of one amplitude
        amp_det += I["k"] * I["B"] * det
        e = I["H"].graph.edge_tensor(dev, "graph").long()
        n_conn = 0
        for bit in (1, 2):
            occ = (I["x"] & bit) != 0
            n_conn += int((occ[:, e[:, 0]] != occ[:, e[:, 1]]).sum().item())
        el_flops += n_conn * mlp
        el_det += n_conn * det
    bf16_net = inst[0]["lp"].precision == "bf16"
    lp_peak = PEAK_BF16 if bf16_net else PEAK_FP32_MFMA
    def mixed_roof(f_bf16, f_f64, ms, what):
        """Two-part roofline: BF16 dense-layer flops at the BF16 matrix peak
        plus FP64 determinant flops at the FP64 peak give the ideal time;
        frac = ideal / measured, peak = total flops / ideal time."""
        t = ms * 1e-3
        ideal = f_bf16 / PEAK_BF16 + f_f64 / PEAK_FP64
        tot = f_bf16 + f_f64
        return {"ms": ms, "bound": "mfma", "achieved": tot / t / 1e12,
                "peak": tot / ideal / 1e12, "unit": "TFLOP/s", "frac": ideal / t,
                "note": (what + ": BF16 dense layers (2 (H^2 + H 2N nf) flops) at the BF16 "
                         "peak plus the FP64 determinant ((2/3) nf^3) at the FP64 peak; "
                         f"{f_bf16 / 1e12:.2f} + {f_f64 / 1e12:.2f} TFLOP per step")}

    kern = {
        "lp_logits": {"ms": comp["lp_logits"], "launches_per_step": n_launch, "bound": "mfma",
                      "achieved": logits_flops / (logits_launch_ms * 1e-3) / 1e12,
                      "peak": lp_peak / 1e12, "unit": "TFLOP/s",
                      "frac": logits_flops / (logits_launch_ms * 1e-3) / lp_peak,
                      "note": (("conv net forward: layers 2-3 on the BF16 matrix cores, layer 1 "
                                "and the head on the vector ALUs" if bf16_net else
                                "conv net forward, FP32 on the vector ALUs") +
                               "; flops of every tap, padding included")},
        "amplitude": mixed_roof(amp_flops, amp_det, comp["amplitude"],
                                "NNBF amplitude of the proposed walkers"),
        "local_energy": mixed_roof(el_flops, el_det, comp["local_energy"],
                                   "NNBF amplitude per connected configuration"),
    }
    kern["per_lattice"] = per_lattice
    dom = max(("lp_logits", "amplitude", "local_energy"), key=lambda k: kern[k]["ms"])
    roof = {kk: kern[dom][kk] for kk in ("bound", "achieved", "peak", "unit", "frac")}
    roof["kernel"] = dom
    roof["traffic"] = None
    out = {
        "metric": "effective samples/s + SR iters/s at 8x8 Hubbard; |dE/site| vs FP64 oracle",
        "value": learned["effective_samples_per_s"],
        "unit": "effective samples/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": learned["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": f"{inst[0]['lp'].precision}-proposal-net/bf16-gemm/f64-det",
        "data": "synthetic (seeded parameters, seeded initial walkers)",
        "config": {"workload": a.workload, "description": w["desc"], "walkers": Btot,
                   "lattices": [I["name"] for I in inst],
                   "walkers_per_lattice": [I["B"] for I in inst],
                   "learned_proposals_per_sample": [I["k"] for I in inst],
                   "plain_proposals_per_sample": [I["E"] for I in inst],
                   "net_channels": C, "train_iters_before": w["train_iters"],
                   "train_every_steps": w["train_every"],
                   "train_walkers_per_lattice": w["train_walkers"],
                   "parallelism": "walkers on 1 GPU"},
        "step": ("one sample per walker of every lattice (learned: |E|/4 proposals; plain: "
                 "one sweep of |E| edge proposals) plus its local energy; a training step "
                 "of the shared net every 10 steps; wavefunction parameters fixed"),
        "learned": learned,
        "plain": plain,
        "ess_ratio_learned_over_plain": (learned["effective_samples_per_s"]
                                         / plain["effective_samples_per_s"]),
        "training": {"seconds": t_train, "heat_bath_loss_first": losses[:len(inst)],
                     "heat_bath_loss_last": losses[len(inst):]},
        "roofline": roof,
        "kernels": kern,
    }
    if not a.no_cpu_baseline:
        import multiprocessing as mp

        from oracle import hubbard as ohub
        from oracle import nnbf as onn
        checks, rates = [], []
        cores = min(os.cpu_count() or 1, 16)
        ctx = mp.get_context("spawn")
        for I in inst:
            params = [p.detach().to(torch.float32).to(torch.float64).cpu().numpy()
                      for p in I["model"].params]
            xs = I["x"][:8].cpu().numpy()
            Eg = I["model"].local_energy(I["x"], I["H"])[:8].cpu().numpy()
            Eo = onn.local_energy(ohub.Hubbard(I["Lx"], I["Ly"], 1.0, 8.0, I["nu"], I["nd"]),
                                  params, xs, 2, w["bf_scale"])
            checks.append({"lattice": I["name"], "mean": float(abs(Eg.mean() - Eo.mean()) / I["N"]),
                           "max_per_sample": float(np.abs(Eg - Eo).max() / I["N"]),
                           "samples": int(len(xs))})
            x_np = I["x"][:1024].cpu().numpy()
            per = max(1, len(x_np) // cores)
            jobs = [("nnbf", I["Lx"], I["Ly"], I["nu"], I["nd"], params, w["bf_scale"],
                     x_np[c * per:(c + 1) * per], 2000 + c, a.cpu_seconds / len(inst), False)
                    for c in range(cores)]
            with ctx.Pool(cores) as pool:
                res = pool.map(_cpu_worker, jobs)
            rates.append(sum(r[0] for r in res) / max(r[1] for r in res))
            log(f"CPU baseline {I['name']}: {rates[-1]:.1f} walker-samples/s")
        out["dE_per_site_vs_fp64_oracle"] = checks
        t_cpu = sum(I["B"] / r for I, r in zip(inst, rates))
        out["cpu_baseline"] = {
            "value": sum(I["B"] / p["tau_int"] for I, p in zip(inst, plain["lattices"])) / t_cpu,
            "unit": "effective samples/s", "cores": cores, "kind": "port",
            "sample": ("FP64 oracle restating the reference's plain sampler (full amplitude per "
                       "proposal, one sweep per sample) and E_loc on each lattice, one process "
                       f"per core, {a.cpu_seconds / len(inst):.0f} s per lattice; rates "
                       "combined over the batch with the GPU plain chain's tau_int"),
            "walker_samples_per_s": dict(zip([I["name"] for I in inst], rates)),
            "cpu_model": _cpu_model(),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
