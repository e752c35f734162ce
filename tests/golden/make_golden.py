"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

* tiny_5x5.npz — the reference's own debug size (5x5 grid, nt=3, 5 realizations,
  1 mode, 8 headings; prel_tests/build_model.cu:944-974) on seeded fields: the seven
  input arrays, the parameters and the full model (S1, S2, val, R).
* obstacles_energy.npz — 7x7, nt=21 with obstacles, a 2x2 terminal subgrid and reward
  method 4 (custom1, alpha 0.3) on a scalar field: inputs and full model.
* c1_digest.json — SHA-256 of the C1 workload's inputs and of its model arrays, with
  nnz and a few sampled rows.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from e2e_gpu_dp_amd.configs import Params, small, workload  # noqa: E402
from e2e_gpu_dp_amd.fields import make_fields  # noqa: E402

PFIELDS = [f for f in Params.__dataclass_fields__]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _save(name, p, f):
    m = oracle.build(p, f)
    np.savez_compressed(
        HERE / name, **{f"in_{k}": v for k, v in f.items()},
        params=np.array(json.dumps({k: getattr(p, k) for k in PFIELDS})),
        S1=m["S1"].astype(np.int32), S2=m["S2"].astype(np.int32), val=m["val"], R=m["R"])
    print(f"{name}: nnz={len(m['S1'])}")


def main():
    w = small(g=5, nt=3, nr=5, nm=1, na=8, seed=2020)
    _save("tiny_5x5.npz", w.params, w.fields())

    p = small(g=7, nt=21, nr=48, nm=2, seed=2021, term_subgrid_size=2, reward_type=4,
              alpha=0.3, num_ac_speeds=2, num_ac_angles=4, F=2.0).params
    f = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, 2021, F=p.F, with_scalar=True,
                    obstacles=[(0, 10, 2, 4, 4, 5), (8, 21, 5, 6, 1, 3)])
    _save("obstacles_energy.npz", p, f)

    w = workload("C1")
    f = w.fields()
    m = oracle.build(w.params, f)
    rows = [0, 1, 777, 123456, len(m["S1"]) - 1]
    dig = {
        "inputs": {k: sha(v) for k, v in f.items()},
        "nnz": int(len(m["S1"])),
        "S1": sha(m["S1"]), "S2": sha(m["S2"]), "val": sha(m["val"]), "R": sha(m["R"]),
        "sample_rows": {str(i): [int(m["S1"][i]), int(m["S2"][i]), float(m["val"][i])] for i in rows},
    }
    (HERE / "c1_digest.json").write_text(json.dumps(dig, indent=1) + "\n")
    print(f"c1_digest.json: nnz={dig['nnz']}")


if __name__ == "__main__":
    main()
