"""CPU checks of the C-ABI library: it loads, exports every symbol include/*.h declares,
and its struct layouts match the Python binding. No compute calls (no GPU here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd import _abi
from e2e_gpu_dp_amd.configs import workload

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(e2e_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    L = _abi.lib()
    names = declared_symbols()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the binding knows all of them
    assert names <= set(_abi.SIGNATURES), names - set(_abi.SIGNATURES)


def test_struct_layouts_match():
    L = _abi.lib()
    for i, s in enumerate(_abi.STRUCTS):
        assert L.e2e_struct_size(i) == C.sizeof(s), s.__name__
    assert L.e2e_struct_size(99) == -1


def test_version_and_device_count():
    L = _abi.lib()
    assert b"gfx950" in L.e2e_version()
    assert isinstance(e2e.device_count(), int)


def test_prob_params_layout():
    """The 32-float vector of src/build_model.cu:1032-1055 (slot meanings per
    src/check_model_output.py:16-35)."""
    p = workload("C2").params.replace(alpha=0.25, is_stationary=1)
    v = e2e.pack_params(p)
    assert v.dtype == np.float32 and v.shape == (32,)
    expect = [p.gsize, p.na, p.nrzns, p.F, p.dt, p.r_outbound, p.r_terminal, p.nmodes, p.i_term,
              p.j_term, p.nt, 1, p.term_subgrid_size, p.reward_type, p.num_ac_speeds,
              p.num_ac_angles, p.dx, p.dy, p.neighb_gsize, 0.25]
    assert np.array_equal(v[:20], np.array(expect, np.float32))
    assert (v[20:] == -9999).all()


def test_action_table():
    """populate_actions (src/build_model.cu:1438-1457): speeds (i+1)F/n, angles
    j*2pi/n, action = speed_idx*n_angles + angle_idx."""
    p = workload("C1").params.replace(num_ac_speeds=3, num_ac_angles=4, F=2.0)
    sp, an = e2e.actions(p)
    assert np.allclose(sp, np.repeat(np.float32([2 / 3, 4 / 3, 2.0]), 4))
    assert np.array_equal(an, np.tile(np.float32([0, np.pi / 2, np.pi, 3 * np.pi / 2]), 3))
    p1 = p.replace(num_ac_speeds=1)
    sp1, _ = e2e.actions(p1)
    assert (sp1 == 2.0).all()


def test_bad_arguments_are_errors_not_crashes():
    L = _abi.lib()
    assert L.e2e_model_info_get(None, None) == 1
    assert L.e2e_model_export_coo(None, None, None, None, None) == 1
    assert L.e2e_vi_solve(None, None, None, None, None) == 1
    assert b"NULL" in L.e2e_last_error() or b"model" in L.e2e_last_error()
    h = C.c_void_p()
    bad = e2e.c_params(workload("C1").params.replace(neighb_gsize=20))
    f = _abi.Fields()
    assert L.e2e_model_build(C.byref(bad), C.byref(f), None, C.byref(h)) == 1
    assert h.value is None


def test_graft_build_entry_compiles_the_oracle():
    """build() compiles the oracle through oracle.compile_lib; oracle.build is the
    oracle's model builder (params, fields), a different function."""
    import inspect

    import oracle
    assert inspect.signature(oracle.compile_lib).parameters.keys() == {"force"}
    assert list(inspect.signature(oracle.build).parameters)[:2] == ["params", "fields"]
    src = (Path(__file__).resolve().parents[1] / "__graft_entry__.py").read_text()
    assert "oracle.compile_lib(" in src
