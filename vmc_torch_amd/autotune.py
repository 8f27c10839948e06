"""Launch-width autotuner for the Slater walker kernels.

The sweep, local-energy and refresh kernels of the Slater path put several
walkers in one workgroup (64 / NFP walkers per wave, `waves` waves per
workgroup) and stage M in LDS once per workgroup.  More waves share one
staging of M; fewer waves give more workgroups to the 256 CUs.  The best
width depends on the walker count, NFP, the state precision and the chip,
so it is measured: `tune_slater` times every candidate width on scratch
copies of a walker batch (the caller's walkers and state are not touched)
and installs the fastest in the library's per-kind table
(vmcx_set_launch_waves), which every later launch reads.

A JSON cache keyed by device name, kernel, NFP and precision keeps the
choice across processes (the last tuned walker count wins): `load_cache`
installs what it holds.

No counterpart in the reference (its sweep is a Python loop of batched torch
calls, GPU/sampler.py:263-388, with fixed launch shapes chosen by torch).
"""
import json
import os

import torch

from . import _lib, kernels

KINDS = {"sweep": 0, "local_energy": 1, "refresh": 2}
CANDIDATES = (1, 2, 4, 8)


def table_key(nfp, fp32):
    """The library's table key for an NFP and a state precision."""
    return 2 * int(nfp) + (1 if fp32 else 0)


def set_waves(kind, nfp, fp32, waves):
    """waves per workgroup for `kind` at (nfp, precision); 0 restores the
    library default."""
    _lib.check(_lib.lib().vmcx_set_launch_waves(KINDS[kind], table_key(nfp, fp32), int(waves)))


def get_waves(kind, nfp, fp32):
    """The installed width (0 = the library default)."""
    return int(_lib.lib().vmcx_get_launch_waves(KINDS[kind], table_key(nfp, fp32)))


def _time(fn, repeats, device):
    fn()                                     # first launch of this shape: untimed
    torch.cuda.synchronize(device)
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(repeats):
        fn()
    e.record()
    torch.cuda.synchronize(device)
    return s.elapsed_time(e) / repeats


def _cache_key(device, kind, nfp, fp32):
    name = torch.cuda.get_device_name(device).replace(" ", "_")
    return f"{name}/{kind}/nfp{nfp}/{'f32' if fp32 else 'f64'}"


def load_cache(path, device=None):
    """Install every width the cache file holds for this device; returns the
    number installed (0 when the file does not exist)."""
    if not path or not os.path.exists(path):
        return 0
    with open(path) as f:
        data = json.load(f)
    if not isinstance(data, dict):
        raise ValueError(f"{path}: not a launch-width cache")
    dev = torch.device("cuda", 0) if device is None else torch.device(device)
    name = torch.cuda.get_device_name(dev).replace(" ", "_")
    n = 0
    for key, waves in data.items():
        parts = key.split("/")
        if len(parts) != 4 or parts[0] != name or parts[1] not in KINDS:
            continue
        nfp, fp32 = int(parts[2][3:]), parts[3] == "f32"
        set_waves(parts[1], nfp, fp32, int(waves))
        n += 1
    return n


def tune_slater(model, fxs, H, kinds=("sweep", "local_energy", "refresh"),
                candidates=CANDIDATES, repeats=5, seed=0, cache=None):
    """Time each candidate width for each kernel kind on copies of the walker
    batch fxs (model: a SlaterDeterminant_GPU; H: its Hamiltonian) and install
    the fastest.  Returns {kind: {"waves": best, "ms": {waves: ms}}}.  cache:
    a JSON path to update with the choices (read by load_cache)."""
    dev = fxs.device
    B, N = fxs.shape
    nf = model.n_fermions
    nfp = kernels.slater_nfp(nf)
    fp32 = model.M.dtype == torch.float32
    M64 = model.M.detach().to(torch.float64).contiguous()
    M32 = model.M.detach().contiguous() if fp32 else None
    x = fxs.clone()
    st = kernels.SlaterWalkerState.empty(B, N, nf, dev, dtype=model.M.dtype)
    kernels.slater_refresh(M64, x, st)
    es = H.graph.edge_tensor(dev, "sweep")
    eg = H.graph.edge_tensor(dev, "graph")
    ops = {
        "sweep": lambda: kernels.slater_sweep(M64, es, x, st, n_sweeps=1, seed=seed,
                                              sweep_offset=0, refresh_every=0, M32=M32),
        "local_energy": lambda: kernels.slater_local_energy(M64, eg, H.t, H.U, x, st, M32=M32),
        "refresh": lambda: kernels.slater_refresh(M64, x, st),
    }
    out = {}
    for kind in kinds:
        if kind not in KINDS:
            raise ValueError(f"unknown kernel kind {kind!r}; one of {sorted(KINDS)}")
        before = get_waves(kind, nfp, fp32)
        ms = {}
        try:
            for w in candidates:
                set_waves(kind, nfp, fp32, w)
                ms[w] = _time(ops[kind], repeats, dev)
        except Exception:
            set_waves(kind, nfp, fp32, before)
            raise
        best = min(ms, key=ms.get)
        set_waves(kind, nfp, fp32, best)
        out[kind] = {"waves": best, "ms": ms}
    if cache:
        data = {}
        if os.path.exists(cache):
            with open(cache) as f:
                data = json.load(f)
        for kind, r in out.items():
            data[_cache_key(dev, kind, nfp, fp32)] = r["waves"]
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return out
