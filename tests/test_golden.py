"""Committed golden fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle and the field recipe must keep reproducing them; the GPU build must match
them bit for bit from the stored inputs."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
import oracle
from e2e_gpu_dp_amd.configs import Params, workload

G = Path(__file__).resolve().parent / "golden"
FIXTURES = ["tiny_5x5.npz", "obstacles_energy.npz"]


def _load(name):
    z = np.load(G / name, allow_pickle=False)
    p = Params(**json.loads(str(z["params"])))
    f = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    return p, f, z


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_golden(name):
    p, f, z = _load(name)
    m = oracle.build(p, f)
    assert np.array_equal(m["S1"], z["S1"]) and np.array_equal(m["S2"], z["S2"])
    assert np.array_equal(m["val"].view(np.uint32), z["val"].view(np.uint32))
    assert np.array_equal(m["R"].view(np.uint32), z["R"].view(np.uint32))


def test_c1_inputs_match_digest():
    d = json.loads((G / "c1_digest.json").read_text())
    f = workload("C1").fields()
    for k, h in d["inputs"].items():
        assert sha(f[k]) == h, k


@pytest.mark.gpu
@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_matches_golden(gpu, name):
    p, f, z = _load(name)
    with e2e.build_model(p, f) as m:
        S1, S2, val, R = m.export_coo()
    assert np.array_equal(S1, z["S1"]) and np.array_equal(S2, z["S2"])
    assert np.array_equal(val.view(np.uint32), z["val"].view(np.uint32))
    assert np.array_equal(R.view(np.uint32), z["R"].view(np.uint32))


@pytest.mark.gpu
def test_gpu_c1_matches_digest(gpu):
    d = json.loads((G / "c1_digest.json").read_text())
    w = workload("C1")
    with e2e.build_model(w.params, w.fields()) as m:
        S1, S2, val, R = m.export_coo()
    assert len(S1) == d["nnz"]
    assert sha(S1) == d["S1"] and sha(S2) == d["S2"] and sha(val) == d["val"] and sha(R) == d["R"]
