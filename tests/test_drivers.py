"""The drop-in executables: build_model / spvi with the reference's files and layout
(src/build_model.cu:926-1314, src/spvi_2.cu:730-796, src/Makefile:10-19)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
import oracle
from e2e_gpu_dp_amd.configs import small
from helpers import rel_err

BIN = Path(e2e.__file__).resolve().parent / "bin"

# parameter file in the syntax of src/input_to_build_model.h (declarations, casts,
# expressions over earlier names, comments)
PARAMS_H = """//Prob type: time, energy1, energy2, energy3
std::string prob_type = "{prob_type}"; //verify num_ac_speeds
std::string prob_name = "synthetic_g{g}";
float alpha_header = 0.05; // 1 for only energy, 0 for only time
int32_t nt = {nt};
float dt = 1;
int32_t gsize = {g};
float dx = 1; float dy = dx;
float x0 = dx/2; //explain later
float y0 = dy/2;
float F = {F};
int num_ac_speeds = {nsp};
int num_ac_angles = 8;
int32_t num_actions = num_ac_speeds*num_ac_angles;
int32_t nrzns = {nr};
int32_t bDimx = 64;
float r_outbound = -100;
float r_terminal = 100;
// int32_t i_term = 50;
int32_t i_term = (int)(0.5*gsize);
int32_t j_term = (int)(0.1*gsize) + 1;
int term_subgrid_size = 2;
float nmodes = {nm};
/* size of neighbourhood grid - odd */
float neighb_gsize = 5;
int32_t is_stationary = 0;
float z = -9999;
"""


def _write_inputs(root: Path, prob_type="time", g=9, nt=6, nr=120, nm=2, nsp=1, F=1.0, seed=61):
    w = small(g=g, nt=nt, nr=nr, nm=nm, seed=seed)
    f = w.fields()
    d = root / "data_input" / f"synthetic_g{g}"
    d.mkdir(parents=True)
    names = {"u": "all_u_mat", "v": "all_v_mat", "ui": "all_ui_mat", "vi": "all_vi_mat",
             "yi": "all_Yi", "s": "all_s_mat", "mask": "obstacle_mask"}
    for k, n in names.items():
        np.save(d / f"{n}.npy", f[k])
    (root / "input_to_build_model.h").write_text(
        PARAMS_H.format(prob_type=prob_type, g=g, nt=nt, nr=nr, nm=nm, nsp=nsp, F=F))
    rt = {"time": 0, "energy1": 1, "energy2": 2, "energy3": 3, "custom1": 4}[prob_type]
    p = w.params.replace(i_term=int(0.5 * g), j_term=int(0.1 * g) + 1, term_subgrid_size=2,
                         neighb_gsize=5, reward_type=rt, num_ac_speeds=nsp, F=F, alpha=0.05)
    return p, f


def _run(args, cwd):
    return subprocess.run([str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=600)


def test_build_model_dry_run_parses_reference_syntax(tmp_path):
    p, _ = _write_inputs(tmp_path, prob_type="custom1")
    r = _run([BIN / "build_model", "30", "--dry-run"], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "temp_modelOp_dirName.txt").read_text().split("\n")
    assert lines[0] == "custom1" and lines[1] == "synthetic_g9"
    assert lines[2] == f"a1x8_i4_j1_ref2"
    assert lines[3] == "0.300000"  # std::to_string(alpha) with alpha = 30/100
    out = tmp_path / lines[4]
    assert lines[4].endswith("/0.300000/")
    pp = np.load(out / "prob_params.npy")
    assert pp.dtype == np.float32 and pp.shape == (32, 1)
    want = e2e.pack_params(p.replace(alpha=np.float32(30) / np.float32(100)))
    assert np.array_equal(pp[:, 0], want)
    dp = np.load(out / "DP_relv_params.npy")
    assert dp.dtype == np.int64 and dp.shape == (2, 1) and list(dp[:, 0]) == [p.ns, p.na]


def test_build_model_rejects_bad_inputs(tmp_path):
    _write_inputs(tmp_path)
    d = tmp_path / "data_input" / "synthetic_g9"
    np.save(d / "all_u_mat.npy", np.zeros((3, 9, 9), np.float32))  # wrong nt
    r = _run([BIN / "build_model", "--dry-run"], tmp_path)
    assert r.returncode == 1 and "all_u_mat" in r.stderr
    np.save(d / "all_u_mat.npy", np.zeros((6, 9, 9), np.float64))  # wrong dtype
    r = _run([BIN / "build_model", "--dry-run"], tmp_path)
    assert r.returncode == 1 and "float32" in r.stderr
    (d / "all_Yi.npy").write_bytes(b"not an npy file")
    np.save(d / "all_u_mat.npy", np.zeros((6, 9, 9), np.float32))
    r = _run([BIN / "build_model", "--dry-run"], tmp_path)
    assert r.returncode == 1 and "all_Yi" in r.stderr


def test_spvi_needs_the_temp_file(tmp_path):
    r = _run([BIN / "spvi"], tmp_path)
    assert r.returncode == 1 and "temp_modelOp_dirName" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("prob_type,int32", [("time", False), ("custom1", True)])
def test_file_pipeline_end_to_end(gpu, tmp_path, prob_type, int32):
    """make run_build_model_and_run_solver: build_model -> (int32 COO) -> spvi."""
    p, f = _write_inputs(tmp_path, prob_type=prob_type, nsp=2 if prob_type == "custom1" else 1,
                         F=2.0 if prob_type == "custom1" else 1.0)
    args = [BIN / "build_model", "40"] + (["--int32"] if int32 else [])
    r = _run(args, tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    lines = (tmp_path / "temp_modelOp_dirName.txt").read_text().split("\n")
    out = tmp_path / lines[4]
    p = p.replace(alpha=np.float32(40) / np.float32(100))
    ref = oracle.build(p, f)
    S1 = np.load(out / "master_cooS1.npy")
    S2 = np.load(out / "master_cooS2.npy")
    assert S1.dtype == (np.int32 if int32 else np.int64) and S1.shape == (len(ref["S1"]), 1)
    assert np.array_equal(S1[:, 0], ref["S1"]) and np.array_equal(S2[:, 0], ref["S2"])
    val = np.load(out / "master_cooVal.npy")[:, 0]
    R = np.load(out / "master_R.npy")[:, 0]
    assert np.array_equal(val.view(np.uint32), ref["val"].view(np.uint32))
    assert np.array_equal(R.view(np.uint32), ref["R"].view(np.uint32))
    if not int32:  # the reference's conversion step, src/convert_model_datatype.py:19-23
        np.save(out / "master_cooS1.npy", S1.astype(np.int32))
        np.save(out / "master_cooS2.npy", S2.astype(np.int32))
    r = _run([BIN / "spvi", "--thresh", "1e-6"], tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    res = tmp_path / "data_solverOutput" / prob_type / "synthetic_g9" / lines[2]
    if prob_type == "custom1":
        res = res / lines[3]
    pol = np.load(res / "policy.npy")
    V = np.load(res / "value_function.npy")
    assert pol.dtype == np.uint32 and pol.shape == (p.ns, 1)
    assert V.dtype == np.float32 and V.shape == (p.ns, 1)
    V64, _ = oracle.vi_backward_f64(p.ns, p.na, p.nc, ref["S1"], ref["S2"], ref["val"], ref["R"])
    assert rel_err(V[:, 0], V64).max() <= 1e-5
