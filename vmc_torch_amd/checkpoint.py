"""Checkpoint and statistics files in the reference's formats, so a run can
resume from (or hand over to) the reference.

GPU runs (GPU/run_scripts/vmcconfig.py:131-224):
  checkpoint_{model}_{step}.pt    torch.save({'model_state_dict', 'fxs'})
                                  (older files: a bare state_dict)
  stats_{model}{suffix}{_from<resume>}.json
                                  {'system', 'Np', 'sample size', 'mean',
                                   'error', 'variance'}, one entry per step
CPU runs (VMC.py:264-299):
  model_params_step{n}.pth        torch.save({'model_structure',
                                  'model_params_vec' (numpy), 'model_state_dict',
                                  'MC_energy_stats'[, 'optimizer_state']})
  energy_stats_start_{start}.json {'Np', 'sample size:', 'mean', 'error',
                                   'variance'}

The models of this package keep the reference's parameter layout
(`params.<i>`, same order and shapes), so state dicts load either way.

GPU checkpoints are read with torch's weights-only loader.  The CPU file
holds numpy objects, which that loader refuses; `read_cpu_checkpoint` decodes
it without unpickling: pickletools walks the opcodes of data.pkl and a small
stack machine rebuilds containers, strings and numbers, and turns the few
object constructors these files use (torch tensors over the archive's
storages, numpy arrays and scalars from their raw bytes, OrderedDict, the
latin-1 bytes codec) into values.  Any other global is refused; nothing named
in the file is imported or called.
"""
import json
import os
import pickletools
import zipfile

import numpy as np
import torch

# ------------------------------------------------------------------ GPU runs


def checkpoint_path(output_dir, model_name, step):
    return os.path.join(output_dir, f"checkpoint_{model_name}_{step}.pt")


def save_checkpoint(model, fxs, output_dir, model_name, step):
    """The reference's checkpoint of step `step` (vmcconfig.py:214-222)."""
    path = checkpoint_path(output_dir, model_name, step)
    torch.save({"model_state_dict": model.state_dict(), "fxs": fxs.detach().cpu()}, path)
    return path


def load_checkpoint(model, output_dir, model_name, resume_step, device, rank=0):
    """vmcconfig.py:131-159: load the parameters of checkpoint `resume_step`
    (> 0) into `model`; both the dict format and a bare state_dict.  Returns
    the saved walkers (on `device`) or None."""
    if resume_step <= 0:
        return None
    path = checkpoint_path(output_dir, model_name, resume_step)
    ckpt = torch.load(path, map_location=device, weights_only=True)
    fxs = None
    if isinstance(ckpt, dict) and "model_state_dict" in ckpt:
        model.load_state_dict(ckpt["model_state_dict"])
        if ckpt.get("fxs") is not None:
            fxs = ckpt["fxs"].to(device)
    else:
        model.load_state_dict(ckpt)
    if hasattr(model, "invalidate"):
        model.invalidate()
    if rank == 0:
        print(f"Loaded checkpoint: {path}")
    return fxs


def make_stats(system_str, n_params, ns_per_rank, world_size):
    """vmcconfig.py:162-171."""
    return {"system": system_str, "Np": n_params, "sample size": ns_per_rank * world_size,
            "mean": [], "error": [], "variance": []}


def make_stats_file(output_dir, model_name, resume_step, suffix=""):
    """vmcconfig.py:174-190."""
    tag = f"_from{resume_step}" if resume_step > 0 else ""
    return os.path.join(output_dir, f"stats_{model_name}{suffix}{tag}.json")


def make_on_step_end(rank, stats, stats_file, output_dir, model_name, model, save_every):
    """vmcconfig.py:193-224: append the step's E/site, error/site and
    variance to the stats file; every `save_every` steps save a checkpoint
    with the walkers (rank 0 only)."""
    def on_step_end(info):
        if rank != 0:
            return
        stats["mean"].append(float(info["energy_per_site"]))
        stats["error"].append(float(info["error_per_site"]))
        stats["variance"].append(float(info["energy_var"]))
        with open(stats_file, "w") as f:
            json.dump(stats, f, indent=4)
        step = info["step"]
        if (step + 1) % save_every == 0:
            save_checkpoint(model, info["fxs"], output_dir, model_name, step + 1)
    return on_step_end


# ------------------------------------------------------------------ CPU runs

_STORAGE_DTYPES = {
    "DoubleStorage": "<f8", "FloatStorage": "<f4", "HalfStorage": "<f2",
    "LongStorage": "<i8", "IntStorage": "<i4", "ShortStorage": "<i2", "CharStorage": "i1",
    "ByteStorage": "u1", "BoolStorage": "?", "BFloat16Storage": "<u2",
}


class _Global:
    def __init__(self, module, name):
        self.module, self.name = module, name

    @property
    def key(self):
        return f"{self.module}.{self.name}".replace("numpy.core.", "numpy._core.")


class _Pending:
    """An object waiting for its BUILD state (numpy array or dtype); BUILD
    completes it in place, so references kept in the memo see the value."""

    def __init__(self, kind, payload=None):
        self.kind, self.payload = kind, payload
        self.done, self.value = False, None


class _Mark:
    pass


def _as_bytes(x):
    return x.encode("latin1") if isinstance(x, str) else bytes(x)


def _finish(v):
    """Resolve completed pending objects (also inside tuples)."""
    if isinstance(v, _Pending):
        if not v.done:
            raise ValueError(f"incomplete {v.kind} in checkpoint")
        return v.value
    if isinstance(v, tuple):
        return tuple(_finish(e) for e in v)
    return v


def _check_tensor_record(n, offset, size, stride):
    """A tensor record must address only its storage's n elements: offset,
    sizes and strides non-negative integers, one stride per dimension, and
    (for a non-empty tensor) offset + sum((size_i - 1) stride_i) < n."""
    ints = all(isinstance(v, int) and not isinstance(v, bool)
               for v in (offset, *size, *stride))
    if not ints or len(size) != len(stride):
        raise ValueError("checkpoint tensor record has a malformed offset, size or stride")
    if offset < 0 or any(d < 0 for d in size) or any(st < 0 for st in stride):
        raise ValueError("checkpoint tensor record has a negative offset, size or stride")
    if not size:
        if offset >= n:
            raise ValueError("checkpoint tensor record lies outside its storage")
        return
    if all(d > 0 for d in size):
        last = offset + sum((d - 1) * st for d, st in zip(size, stride))
        if last >= n:
            raise ValueError("checkpoint tensor record lies outside its storage")
    elif offset > n:
        raise ValueError("checkpoint tensor record lies outside its storage")


def _call(fn, args, storages):
    args = _finish(args)
    k = fn.key
    if k == "collections.OrderedDict":
        return {}
    if k == "_codecs.encode":
        return _as_bytes(args[0])
    if k == "torch._utils._rebuild_tensor_v2":
        storage, offset, size, stride = args[0], args[1], tuple(args[2]), tuple(args[3])
        flat = storage
        _check_tensor_record(len(flat), offset, size, stride)
        if not size:
            return flat[offset:offset + 1].reshape(()).copy()
        isz = flat.dtype.itemsize
        view = np.lib.stride_tricks.as_strided(flat[offset:], shape=size,
                                               strides=[s * isz for s in stride])
        return view.copy()
    if k == "numpy._core.multiarray._reconstruct":
        return _Pending("ndarray")
    if k == "numpy.dtype":
        return _Pending("dtype", str(args[0]))
    if k == "numpy._core.multiarray.scalar":
        return np.frombuffer(_as_bytes(args[1]), dtype=args[0])[0]
    raise ValueError(f"checkpoint uses an unsupported object: {fn.module}.{fn.name}")


def _build(obj, state):
    state = _finish(state)
    if isinstance(obj, _Pending) and not obj.done and obj.kind == "dtype":
        order = state[1] if isinstance(state, tuple) and len(state) > 1 else "="
        dt = np.dtype(obj.payload)
        obj.value, obj.done = (dt.newbyteorder(order) if order in ("<", ">") else dt), True
        return obj
    if isinstance(obj, _Pending) and not obj.done and obj.kind == "ndarray":
        _, shape, dt, fortran, raw = state[:5]
        a = np.frombuffer(_as_bytes(raw), dtype=dt).copy()
        obj.value = a.reshape(tuple(shape), order="F" if fortran else "C")
        obj.done = True
        return obj
    if isinstance(obj, dict) and isinstance(state, dict):
        return obj      # instance attributes of an OrderedDict (state_dict._metadata)
    raise ValueError("checkpoint uses an unsupported BUILD")


def _decode(pkl, read_storage):
    stack, memo, storages = [], {}, {}

    def pop_mark():
        i = len(stack) - 1
        while not isinstance(stack[i], _Mark):
            i -= 1
        items = stack[i + 1:]
        del stack[i:]
        return items
    for op, arg, _ in pickletools.genops(pkl):
        n = op.name
        if n in ("PROTO", "FRAME"):
            continue
        if n == "STOP":
            return _finish(stack.pop())
        if n == "MARK":
            stack.append(_Mark())
        elif n in ("EMPTY_DICT",):
            stack.append({})
        elif n in ("EMPTY_LIST",):
            stack.append([])
        elif n == "EMPTY_TUPLE":
            stack.append(())
        elif n in ("BINUNICODE", "SHORT_BINUNICODE", "BINUNICODE8", "UNICODE",
                   "BINSTRING", "SHORT_BINSTRING", "STRING"):
            stack.append(arg)
        elif n in ("BINBYTES", "SHORT_BINBYTES", "BINBYTES8"):
            stack.append(bytes(arg))
        elif n in ("BININT", "BININT1", "BININT2", "LONG1", "LONG4", "INT", "LONG",
                   "BINFLOAT", "FLOAT"):
            stack.append(arg)
        elif n == "NONE":
            stack.append(None)
        elif n == "NEWTRUE":
            stack.append(True)
        elif n == "NEWFALSE":
            stack.append(False)
        elif n in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[arg] = stack[-1]
        elif n == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif n in ("BINGET", "LONG_BINGET", "GET"):
            stack.append(memo[arg])
        elif n == "TUPLE":
            stack.append(tuple(_finish(e) for e in pop_mark()))
        elif n in ("TUPLE1", "TUPLE2", "TUPLE3"):
            k = {"TUPLE1": 1, "TUPLE2": 2, "TUPLE3": 3}[n]
            items = tuple(_finish(e) for e in stack[-k:])
            del stack[-k:]
            stack.append(items)
        elif n == "LIST":
            stack.append([_finish(e) for e in pop_mark()])
        elif n == "DICT":
            items = [_finish(e) for e in pop_mark()]
            stack.append(dict(zip(items[0::2], items[1::2])))
        elif n == "APPEND":
            v = stack.pop()
            stack[-1].append(_finish(v))
        elif n == "APPENDS":
            items = pop_mark()
            stack[-1].extend(_finish(e) for e in items)
        elif n == "SETITEM":
            v = stack.pop()
            k = stack.pop()
            stack[-1][k] = _finish(v)
        elif n == "SETITEMS":
            items = pop_mark()
            for k, v in zip(items[0::2], items[1::2]):
                stack[-1][k] = _finish(v)
        elif n == "GLOBAL":
            mod, name = arg.split(" ", 1) if " " in arg else arg.split("\n", 1)
            stack.append(_Global(mod, name))
        elif n == "STACK_GLOBAL":
            name = stack.pop()
            mod = stack.pop()
            stack.append(_Global(mod, name))
        elif n == "REDUCE":
            args = stack.pop()
            fn = stack.pop()
            if not isinstance(fn, _Global):
                raise ValueError("checkpoint calls a non-global")
            stack.append(_call(fn, args, storages))
        elif n == "NEWOBJ":
            args = stack.pop()
            cls = stack.pop()
            stack.append(_call(cls, args, storages))
        elif n == "BUILD":
            state = stack.pop()
            stack[-1] = _build(stack[-1], state)
        elif n in ("BINPERSID", "PERSID"):
            pid = _finish(stack.pop())
            if not (isinstance(pid, tuple) and pid and pid[0] == "storage"):
                raise ValueError("checkpoint has an unknown persistent id")
            stype, key = pid[1], str(pid[2])
            name = stype.name if isinstance(stype, _Global) else str(stype)
            if name not in _STORAGE_DTYPES:
                raise ValueError(f"checkpoint storage type {name} is not supported")
            if key not in storages:
                storages[key] = np.frombuffer(read_storage(key), dtype=_STORAGE_DTYPES[name])
            stack.append(storages[key])
        else:
            raise ValueError(f"checkpoint uses pickle opcode {n}, not supported")
    raise ValueError("checkpoint pickle has no STOP")


def read_cpu_checkpoint(path):
    """Contents of a reference CPU checkpoint (model_params_step{n}.pth) as
    plain Python / numpy values: tensors become numpy arrays (BFloat16 as
    their raw uint16 bits)."""
    with zipfile.ZipFile(path) as z:
        names = z.namelist()
        pk = [n for n in names if n.endswith("/data.pkl") or n == "data.pkl"]
        if len(pk) != 1:
            raise ValueError(f"{path}: not a torch zip checkpoint")
        prefix = pk[0][: -len("data.pkl")]
        pkl = z.read(pk[0])
        return _decode(pkl, lambda key: z.read(f"{prefix}data/{key}"))


def load_cpu_params(model, path):
    """Load the parameters of a reference CPU checkpoint into `model`;
    returns the decoded file.  Names that match the model's state dict load
    directly.  A one-tensor CPU model into a one-parameter model (the CPU
    Slater determinant, key 'M') loads by position; the CPU Slater matrix has
    more rows than spin-orbitals (model.py:7416-7457 sizes it by the Hilbert
    space) and only its first 2N rows are ever indexed, so those are taken."""
    d = read_cpu_checkpoint(path)
    sd = {k: v for k, v in d["model_state_dict"].items() if isinstance(v, np.ndarray)}
    own = model.state_dict()
    new = {}
    if all(k in own for k in sd):
        for k, v in sd.items():
            new[k] = torch.as_tensor(np.ascontiguousarray(v), dtype=own[k].dtype)
    elif len(sd) == 1 and len([k for k in own if k.startswith("params.")]) == 1:
        (v,) = sd.values()
        want = own["params.0"].shape
        if v.ndim != 2 or v.shape[1] != want[1] or v.shape[0] < want[0]:
            raise ValueError(f"checkpoint tensor {v.shape} does not fit {tuple(want)}")
        new["params.0"] = torch.as_tensor(np.ascontiguousarray(v[: want[0]]),
                                          dtype=own["params.0"].dtype)
    else:
        raise KeyError(f"checkpoint parameters {sorted(sd)} do not match "
                       f"{type(model).__name__} ({sorted(own)})")
    model.load_state_dict(new, strict=False)
    if hasattr(model, "invalidate"):
        model.invalidate()
    return d


def cpu_params_path(output_dir, step):
    return os.path.join(output_dir, f"model_params_step{step}.pth")


def energy_stats_path(output_dir, start):
    return os.path.join(output_dir, f"energy_stats_start_{start}.json")


def save_cpu_checkpoint(model, output_dir, step, mc_energy_stats, model_structure=None):
    """A file in the reference's CPU format (VMC.py:268-295): model
    structure, flat parameter vector (numpy), state dict and the step's MC
    energy statistics."""
    vec = torch.cat([p.detach().reshape(-1).cpu() for p in model.parameters()]).numpy()
    data = {"model_structure": model_structure or {type(model).__name__: {}},
            "model_params_vec": vec,
            "model_state_dict": {k: v.detach().cpu() for k, v in model.state_dict().items()},
            "MC_energy_stats": mc_energy_stats}
    path = cpu_params_path(output_dir, step)
    torch.save(data, path)
    return path


def make_cpu_stats(n_params, n_samples):
    """The CPU run's statistics dict (VMC.py:135)."""
    return {"Np": n_params, "sample size:": n_samples, "mean": [], "error": [],
            "variance": []}


def make_cpu_on_step_end(output_dir, start, model, stats, n_sites, save_every=1):
    """VMC.py:264-299 as an on_step_end callback: after step s the updated
    parameters go to model_params_step{s+1}.pth with the step's MC energy
    statistics (total energy, not per site, as the CPU driver logs them),
    and energy_stats_start_{start}.json is rewritten."""
    def on_step_end(info):
        mc = {"mean": float(info["energy_mean"]), "variance": float(info["energy_var"]),
              "error": float(info["error_per_site"]) * n_sites}
        stats["mean"].append(mc["mean"])
        stats["error"].append(mc["error"])
        stats["variance"].append(mc["variance"])
        step = info["step"]
        if step % save_every == 0:
            save_cpu_checkpoint(model, output_dir, step + 1, mc)
        with open(energy_stats_path(output_dir, start), "w") as f:
            json.dump(stats, f, indent=4)
    return on_step_end
