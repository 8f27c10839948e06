"""Extract the reference's 4x2 Slater run fixtures into a small JSON file.

Reads, without unpickling, the data files shipped with the reference:
  examples/example_data/4x2/t=1.0_U=8.0/N=8/SlaterDeterminant/
    energy_stats_start_0.json         per-step VMC mean / error / variance
    model_params_step{1,26,50}.pth    zip archives; member <name>/data/0 is the
                                      raw little-endian float64 storage of M
                                      (4900 x 8, row-major; only rows 0..15
                                      are ever indexed by the model)
Writes tests/golden/ref_4x2_slater.json.  Run in a container that has the
reference checkout at /root/reference.
"""
import json
import os
import zipfile

import numpy as np

REF = "/root/reference/examples/example_data/4x2/t=1.0_U=8.0/N=8/SlaterDeterminant"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_4x2_slater.json")


def main():
    stats = json.load(open(os.path.join(REF, "energy_stats_start_0.json")))
    ckpt = {}
    for s in (1, 26, 50):
        z = zipfile.ZipFile(os.path.join(REF, f"model_params_step{s}.pth"))
        raw = z.read(f"model_params_step{s}/data/0")
        M = np.frombuffer(raw, dtype="<f8").reshape(4900, 8)
        ckpt[str(s)] = {"M16": M[:16].tolist(),
                        "inert_rows_norm": float(np.linalg.norm(M[16:]))}
    out = {
        "source": "examples/example_data/4x2/t=1.0_U=8.0/N=8/SlaterDeterminant",
        "system": {"Lx": 4, "Ly": 2, "t": 1.0, "U": 8.0, "n_up": 4, "n_dn": 4},
        "Np": stats["Np"], "n_samples": stats["sample size:"],
        "mean": stats["mean"], "error": stats["error"], "variance": stats["variance"],
        "checkpoints": ckpt,
        "note": ("checkpoint step s holds the parameters used at loop step s; "
                 "its sampled energy is mean[s] (step 50 was never sampled)"),
    }
    json.dump(out, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
