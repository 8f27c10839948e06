"""Checkpoint and statistics files in the reference's formats (CPU only).

GPU-run files (vmcconfig.py:131-224) round-trip through save / load with the
walkers; CPU-run files (VMC.py:264-299) are decoded without unpickling and
agree with what torch.save wrote; the decoder refuses any object outside the
handful these files use without calling it; the reference's shipped 4x2
checkpoints (when the reference checkout is present) decode to the orbital
matrix of the committed fixture and the logged energies."""
import json
import os

import numpy as np
import pytest
import torch

from vmc_torch_amd import NNBF_GPU, SlaterDeterminant_GPU
from vmc_torch_amd import checkpoint as ck

REF = "/root/reference/examples/example_data/4x2/t=1.0_U=8.0/N=8/SlaterDeterminant"
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_4x2_slater.json")


def test_gpu_checkpoint_roundtrip(tmp_path):
    g = torch.Generator().manual_seed(0)
    m = NNBF_GPU(8, 8, hidden_dim=16, generator=g)
    fxs = torch.randint(0, 4, (32, 8))
    ck.save_checkpoint(m, fxs, str(tmp_path), "NNBF", 7)
    m2 = NNBF_GPU(8, 8, hidden_dim=16, generator=torch.Generator().manual_seed(1))
    got = ck.load_checkpoint(m2, str(tmp_path), "NNBF", 7, "cpu")
    assert torch.equal(got, fxs)
    for a, b in zip(m.params, m2.params):
        assert torch.equal(a, b)
    # a bare state_dict (the older format) loads too, and returns no walkers
    torch.save(m.state_dict(), ck.checkpoint_path(str(tmp_path), "NNBF", 8))
    assert ck.load_checkpoint(m2, str(tmp_path), "NNBF", 8, "cpu") is None
    assert ck.load_checkpoint(m2, str(tmp_path), "NNBF", 0, "cpu") is None


def test_gpu_stats_and_callback(tmp_path):
    m = SlaterDeterminant_GPU(16, 8, generator=torch.Generator().manual_seed(2))
    stats = ck.make_stats("4x2 Hubbard", 128, 512, 2)
    path = ck.make_stats_file(str(tmp_path), "Slater", 0)
    assert path.endswith("stats_Slater.json")
    assert ck.make_stats_file(str(tmp_path), "Slater", 5, "_reuse").endswith(
        "stats_Slater_reuse_from5.json")
    cb = ck.make_on_step_end(0, stats, path, str(tmp_path), "Slater", m, save_every=2)
    fxs = torch.zeros(4, 8, dtype=torch.int64)
    for s in range(4):
        cb({"step": s, "energy_per_site": -0.5 - s, "error_per_site": 0.01,
            "energy_var": 2.0 + s, "fxs": fxs})
    got = json.load(open(path))
    assert got == {"system": "4x2 Hubbard", "Np": 128, "sample size": 1024,
                   "mean": [-0.5, -1.5, -2.5, -3.5], "error": [0.01] * 4,
                   "variance": [2.0, 3.0, 4.0, 5.0]}
    assert sorted(os.listdir(tmp_path)) == ["checkpoint_Slater_2.pt", "checkpoint_Slater_4.pt",
                                            "stats_Slater.json"]


def test_cpu_checkpoint_decodes_what_torch_saved(tmp_path):
    m = NNBF_GPU(8, 8, hidden_dim=16, generator=torch.Generator().manual_seed(3))
    mc = {"mean": np.float64(-3.25), "variance": np.float64(1.5), "error": 0.125}
    path = ck.save_cpu_checkpoint(m, str(tmp_path), 4, mc,
                                  {"NNBF": {"N_sites": 8, "hidden": 16}})
    d = ck.read_cpu_checkpoint(path)
    assert d["model_structure"] == {"NNBF": {"N_sites": 8, "hidden": 16}}
    vec = torch.cat([p.detach().reshape(-1) for p in m.params]).numpy()
    assert d["model_params_vec"].dtype == vec.dtype
    assert np.array_equal(d["model_params_vec"], vec)
    for k, v in m.state_dict().items():
        assert np.array_equal(d["model_state_dict"][k], v.numpy())
    assert d["MC_energy_stats"] == {"mean": -3.25, "variance": 1.5, "error": 0.125}
    m2 = NNBF_GPU(8, 8, hidden_dim=16, generator=torch.Generator().manual_seed(4))
    ck.load_cpu_params(m2, path)
    for a, b in zip(m.params, m2.params):
        assert torch.equal(a, b)
    # non-contiguous and strided tensors, other dtypes, nested containers
    t = torch.arange(24, dtype=torch.float32).reshape(4, 6)[:, ::2]
    obj = {"t": t, "i": torch.arange(5), "b": torch.tensor([True, False]),
           "lst": [1, 2.5, "x", (3, None)], "arr": np.arange(6, dtype=np.int32).reshape(2, 3).T}
    torch.save(obj, str(tmp_path / "misc.pth"))
    got = ck.read_cpu_checkpoint(str(tmp_path / "misc.pth"))
    assert np.array_equal(got["t"], t.numpy()) and got["t"].dtype == np.float32
    assert np.array_equal(got["i"], np.arange(5)) and got["b"].tolist() == [True, False]
    assert got["lst"] == [1, 2.5, "x", (3, None)]
    assert np.array_equal(got["arr"], obj["arr"])


def test_cpu_callback_writes_reference_layout(tmp_path):
    m = SlaterDeterminant_GPU(16, 8, generator=torch.Generator().manual_seed(5))
    stats = ck.make_cpu_stats(128, 1024)
    cb = ck.make_cpu_on_step_end(str(tmp_path), 0, m, stats, n_sites=8)
    for s in range(2):
        cb({"step": s, "energy_mean": -4.0 * (s + 1), "energy_per_site": -0.5 * (s + 1),
            "error_per_site": 0.01, "energy_var": 3.0})
    got = json.load(open(tmp_path / "energy_stats_start_0.json"))
    assert got == {"Np": 128, "sample size:": 1024, "mean": [-4.0, -8.0],
                   "error": [0.08, 0.08], "variance": [3.0, 3.0]}
    d = ck.read_cpu_checkpoint(str(tmp_path / "model_params_step2.pth"))
    assert d["MC_energy_stats"] == {"mean": -8.0, "variance": 3.0, "error": 0.08}
    assert np.array_equal(d["model_state_dict"]["params.0"], m.M.detach().numpy())


CALLED = []


def _side_effect(*args):
    CALLED.append(args)
    return 0


class _Hostile:
    def __reduce__(self):
        return (_side_effect, ("ran",))


def test_cpu_checkpoint_refuses_other_objects(tmp_path):
    torch.save({"x": _Hostile()}, str(tmp_path / "bad.pth"))
    with pytest.raises(ValueError, match="unsupported object"):
        ck.read_cpu_checkpoint(str(tmp_path / "bad.pth"))
    assert CALLED == []


def _tensor_record_file(path, n_storage, offset, size, stride):
    """A torch-format zip whose data.pkl (written opcode by opcode) holds one
    _rebuild_tensor_v2 record over an FP64 storage of n_storage elements."""
    import struct
    import zipfile

    def uni(t):
        b = t.encode()
        return b"X" + struct.pack("<I", len(b)) + b

    def i32(v):
        return b"J" + struct.pack("<i", v)

    def tup(vals):
        return b"(" + b"".join(i32(v) for v in vals) + b"t"
    pkl = (b"\x80\x02ctorch._utils\n_rebuild_tensor_v2\n("
           + b"(" + uni("storage") + b"ctorch\nDoubleStorage\n" + uni("0") + uni("cpu")
           + i32(n_storage) + b"tQ"
           + i32(offset) + tup(size) + tup(stride) + b"\x89" + b"}" + b"tR.")
    with zipfile.ZipFile(path, "w") as z:
        z.writestr("archive/data.pkl", pkl)
        z.writestr("archive/data/0", np.arange(n_storage, dtype=np.float64).tobytes())
    return str(path)


def test_cpu_checkpoint_tensor_record_bounds(tmp_path):
    """Tensor records are checked against their storage before any view is
    built: a record past the end, a negative offset or stride, or a stride
    count that does not match the sizes is refused."""
    ok = _tensor_record_file(tmp_path / "ok.pth", 4, 2, (2,), (1,))
    assert np.array_equal(ck.read_cpu_checkpoint(ok), [2.0, 3.0])
    ok2 = _tensor_record_file(tmp_path / "ok2.pth", 6, 1, (2, 2), (3, 1))
    assert np.array_equal(ck.read_cpu_checkpoint(ok2), [[1.0, 2.0], [4.0, 5.0]])
    for i, (off, size, stride) in enumerate([(2, (8,), (1,)), (-1, (2,), (1,)),
                                             (0, (2,), (-1,)), (0, (2, 2), (3,)),
                                             (4, (), ()), (0, (3,), (2,))]):
        bad = _tensor_record_file(tmp_path / f"bad{i}.pth", 4, off, size, stride)
        with pytest.raises(ValueError, match="tensor record"):
            ck.read_cpu_checkpoint(bad)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
def test_reference_cpu_checkpoints():
    gold = json.load(open(GOLD))
    for s in (1, 26, 50):
        m = SlaterDeterminant_GPU(16, 8)
        d = ck.load_cpu_params(m, os.path.join(REF, f"model_params_step{s}.pth"))
        assert d["model_structure"] == {"SlaterDeterminant": {"N_orbitals": 4900,
                                                              "N_fermions": 8}}
        assert np.array_equal(m.M.detach().numpy(), np.array(gold["checkpoints"][str(s)]["M16"]))
        assert np.array_equal(d["model_params_vec"], d["model_state_dict"]["M"].reshape(-1))
        # the file of step s carries the statistics of loop step s - 1
        assert float(d["MC_energy_stats"]["mean"]) == gold["mean"][s - 1]


@pytest.mark.gpu
def test_resume_on_device(tmp_path, device):
    """A checkpoint written from a device model loads into a fresh device
    model whose cached walker state is stale: after the load its amplitudes
    of the saved walkers equal the original's, bit for bit."""
    from helpers import random_configs
    g = torch.Generator().manual_seed(6)
    m = NNBF_GPU(16, 14, hidden_dim=32, generator=g).to(device)
    fxs = torch.tensor(random_configs(np.random.default_rng(7), 64, 16, 7, 7), device=device)
    s0, l0 = m.forward_log(fxs)
    ck.save_checkpoint(m, fxs, str(tmp_path), "NNBF", 3)
    m2 = NNBF_GPU(16, 14, hidden_dim=32, generator=torch.Generator().manual_seed(8)).to(device)
    m2.forward_log(fxs)                      # builds state for the old parameters
    got = ck.load_checkpoint(m2, str(tmp_path), "NNBF", 3, device)
    assert got.device == fxs.device and torch.equal(got, fxs)
    s1, l1 = m2.forward_log(got)
    assert torch.equal(s0, s1) and torch.equal(l0, l1)
