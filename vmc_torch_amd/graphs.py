"""HIP-graph capture of the sampling loop.

A captured graph replays a fixed sequence of libvmcx launches on fixed
buffers with one host call: the sweeps (or learned proposals), the local
energies and the log-derivatives of a walker batch, or the proposals of
several lattice instances at once.  Only libvmcx kernels are captured (no
library call that may allocate or record memset nodes); the Philox counters
of the draws live in device memory and the graph advances them
(vmcx_counter_add), so replay k draws the same uniforms as the eager call
with offset offset0 + k * (draws per replay).  Capture uses torch's graph
object (torch.cuda.CUDAGraph, a hipGraph on ROCm) on a walker batch whose
buffers stay in place.

Replaces: the reference's per-edge Python loop of launches
(GPU/sampler.py:263-388) and its torch.compile / export fusion of the
amplitude (GPU/models/_base.py:201-262).
"""
import torch

from . import kernels


def _capture(body, device):
    """One eager-free capture of `body` into a new graph (torch's capture
    stream becomes the current stream, so every libvmcx launch is recorded)."""
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize(device)
    with torch.cuda.graph(g):
        body()
    torch.cuda.synchronize(device)
    return g


class SlaterSampleGraph:
    """One sample of the Slater path for the walker batch `fxs` as one
    graph: n_sweeps Metropolis sweeps (exact refresh of the walkers due, every
    sweep), the local energies and the log-derivatives.  model: a
    SlaterDeterminant_GPU whose parameters stay fixed between replays (call
    `update()` after they change: it refreshes the walker state and
    recaptures).  replay() returns (E_loc [B], O [B, Np]), static tensors
    overwritten by the next replay."""

    def __init__(self, model, fxs, H, n_sweeps=1, seed=0, sweep_offset=0, refresh_every=64,
                 o_dtype=torch.float64):
        if not getattr(model, "graph_safe", False):
            raise TypeError(f"{type(model).__name__} launches library calls: not capturable")
        self.model, self.fxs, self.H = model, fxs, H
        self.n_sweeps, self.seed, self.refresh_every = int(n_sweeps), int(seed), refresh_every
        dev = fxs.device
        self.counter = torch.tensor([int(sweep_offset)], dtype=torch.int64, device=dev)
        B, N = fxs.shape
        self.E = torch.empty(B, dtype=torch.float64, device=dev)
        self.O = torch.empty(B, model.n_params, dtype=o_dtype, device=dev)
        self.e_sweep = H.graph.edge_tensor(dev, "sweep")
        self.e_graph = H.graph.edge_tensor(dev, "graph")
        self._warm()
        self.update()

    def _body(self, state, fxs, E, O, counter):
        m = self.model
        M, M32 = m._m_f64(), (m.M.detach() if m.fp32 else None)
        kernels.slater_sweep(M, self.e_sweep, fxs, state, n_sweeps=self.n_sweeps, seed=self.seed,
                             sweep_offset=0, refresh_every=self.refresh_every, M32=M32,
                             offset_dev=counter)
        kernels.counter_add(counter, self.n_sweeps)
        kernels.slater_local_energy(M, self.e_graph, self.H.t, self.H.U, fxs, state, out=E,
                                    M32=M32)
        kernels.slater_logderiv(state, m.n_sites, m.n_fermions, dtype=O.dtype, out=O)

    def _warm(self):
        """Load every kernel once on scratch copies (no launch inside the
        capture may be a first launch)."""
        m = self.model
        st = kernels.SlaterWalkerState.empty(self.fxs.shape[0], m.n_sites, m.n_fermions,
                                             self.fxs.device, dtype=m.M.dtype)
        x = self.fxs.clone()
        with m.frozen():
            kernels.slater_refresh(m._m_f64(), x, st)
            self._body(st, x, torch.empty_like(self.E), torch.empty_like(self.O),
                       self.counter.clone())
        torch.cuda.synchronize(self.fxs.device)

    def update(self):
        """(Re)build the walker state from the current parameters and capture."""
        m = self.model
        with m.frozen():
            m.walker_state(self.fxs, refresh=True)
            self.state = m._state
            self.graph = _capture(lambda: self._body(self.state, self.fxs, self.E, self.O,
                                                     self.counter), self.fxs.device)

    def replay(self):
        self.graph.replay()
        return self.E, self.O


class LearnedProposalGraph:
    """n_steps learned proposals for one or several (sampler, model, walkers)
    instances -- the lattices of an uneven batch, each with its own model and
    walker batch and possibly one shared net -- captured as ONE graph: per
    instance and step the net's proposal (vmcx_lp_propose_given), the
    model's amplitudes of the proposed configurations (its libvmcx forward;
    for NNBF the moved-configuration forward from the walkers' first-layer
    state, which each replay first recomputes from the configurations, as
    an eager sweeps() call does) and the Metropolis-Hastings test
    (vmcx_lp_accept, vmcx_lp_commit_logits).
    steps: proposals per replay for each instance (an int or one per
    instance).  The proposal counter of each instance is device-resident;
    replay k of instance i draws what LearnedProposalSamplerGPU.sweeps
    draws from step_counter = start_i + k * steps_i.  The models' parameters
    and the nets stay fixed while the graph is in use (call `update()` after
    a change).  The graph owns the samplers' draw counters while in use
    (sync_counters() hands them back).  After replay(): sign / logabs /
    n_accept per instance."""

    def __init__(self, instances, steps):
        self.inst = []
        steps = [steps] * len(instances) if isinstance(steps, int) else list(steps)
        if len(steps) != len(instances):
            raise ValueError("one step count per instance")
        for (sampler, model, fxs), k in zip(instances, steps):
            if getattr(sampler, "moves_per_proposal", 1) != 1:
                raise NotImplementedError("captured sweeps make one move per proposal")
            if not getattr(model, "graph_safe", False):
                raise TypeError(f"{type(model).__name__} (this precision) launches library "
                                "calls: not capturable")
            dev = fxs.device
            if sampler.params.device != dev:
                sampler.params = sampler.params.to(dev)
            B, N = fxs.shape
            M = kernels.lp_n_moves(sampler.Lx, sampler.Ly)
            I = dict(sampler=sampler, model=model, fxs=fxs, k=int(k),
                     counter=torch.tensor([sampler.step_counter], dtype=torch.int64, device=dev),
                     sign=torch.empty(B, dtype=torch.float64, device=dev),
                     logabs=torch.empty(B, dtype=torch.float64, device=dev),
                     n_acc=torch.zeros(B, dtype=torch.int64, device=dev),
                     lx=torch.empty(B, M, dtype=torch.float64, device=dev),
                     acc=torch.empty(B, dtype=torch.uint8, device=dev),
                     moved=hasattr(model, "forward_log_moved"),
                     bufs=(torch.empty(B, dtype=torch.int32, device=dev),
                           torch.empty(B, dtype=torch.float64, device=dev),
                           torch.empty(B, N, dtype=torch.int64, device=dev),
                           torch.empty(B, dtype=torch.float64, device=dev),
                           torch.empty(B, M, dtype=torch.float64, device=dev)))
            if I["moved"]:
                # first-layer state of the walkers and of the proposals
                I["z1"] = torch.empty(B, model.hidden, dtype=model.z_dtype, device=dev)
                I["mout"] = (torch.empty(B, dtype=torch.float64, device=dev),
                             torch.empty(B, dtype=torch.float64, device=dev),
                             torch.empty(B, model.hidden, dtype=model.z_dtype, device=dev))
            self.inst.append(I)
        self.device = self.inst[0]["fxs"].device
        self._warm()
        self.update()

    def _steps(self, I, fxs, sign, logabs, lx, n_acc, counter, z1=None):
        s, m = I["sampler"], I["model"]
        move, lqf, prop, lqr, lp = I["bufs"]
        for _ in range(I["k"]):
            kernels.lp_propose_given(s.params, s.C, s.Lx, s.Ly, fxs, lx, s.seed, 0,
                                     precision=s.precision, step_dev=counter,
                                     out=(move, lqf, prop, lqr, lp))
            if I["moved"]:
                ps, pl, z1p = m.forward_log_moved(fxs, prop, z1, out=I["mout"])
            else:
                ps, pl = m.forward_log(prop)
            kernels.lp_accept_commit(fxs, sign, logabs, move, prop, ps, pl, lqf, lqr, s.seed, 0,
                                     lp, lx, I["acc"], n_accept=n_acc, step_dev=counter,
                                     z1p=z1p if I["moved"] else None,
                                     z1=z1 if I["moved"] else None)
            kernels.counter_add(counter, 1)

    def _warm(self):
        for I in self.inst:
            m, s = I["model"], I["sampler"]
            x = I["fxs"].clone()
            with m.frozen():
                z1 = None
                if I["moved"]:
                    sg, la, z1 = m.first_layer(x)
                else:
                    sg, la = m.forward_log(x)
                lx = kernels.lp_logits(s.params, s.C, s.Lx, s.Ly, x, precision=s.precision)
                self._steps(I, x, sg.clone(), la.clone(), lx, torch.zeros_like(I["n_acc"]),
                            I["counter"].clone(), z1)
        torch.cuda.synchronize(self.device)

    def update(self):
        """Amplitudes and logits of the current walkers, then the capture."""
        for I in self.inst:
            m, s = I["model"], I["sampler"]
            with m.frozen():
                sg, la = m.forward_log(I["fxs"])
                I["sign"].copy_(sg)
                I["logabs"].copy_(la)
                I["lx"].copy_(kernels.lp_logits(s.params, s.C, s.Lx, s.Ly, I["fxs"],
                                                precision=s.precision))

        def body():
            for I in self.inst:
                if I["moved"]:
                    # the first-layer state (and the amplitudes) of the walkers
                    # from their configurations, as every eager sweeps() call
                    # starts
                    I["model"].first_layer(I["fxs"], out=(I["sign"], I["logabs"], I["z1"]))
                self._steps(I, I["fxs"], I["sign"], I["logabs"], I["lx"], I["n_acc"],
                            I["counter"], I.get("z1"))
        for I in self.inst:
            I["model"]._frozen = True
        try:
            self.graph = _capture(body, self.device)
        finally:
            for I in self.inst:
                I["model"]._frozen = False
        for I in self.inst:
            if hasattr(I["model"], "invalidate"):
                I["model"].invalidate()

    def sync_counters(self):
        """Hand the draw counters back to the samplers (one host read), so
        eager sweeps continue where the graph stopped."""
        for I in self.inst:
            I["sampler"].step_counter = int(I["counter"].item())

    def replay(self):
        """One replay: steps_i proposals for every instance."""
        self.graph.replay()
        for I in self.inst:
            if hasattr(I["model"], "invalidate"):
                I["model"].invalidate()   # the cached walker states predate the moves
        return [(I["sign"], I["logabs"]) for I in self.inst]
