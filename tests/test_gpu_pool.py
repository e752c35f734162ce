"""Device blocks kept mapped after a model closes (devpool.cpp) serve the next models:
builds and solves in reused memory give the same bits as in fresh memory, and the idle
blocks go back to the runtime on request."""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import workload

pytestmark = pytest.mark.gpu


def test_reused_device_memory_gives_the_same_models_and_solves(gpu):
    w = workload("C1")
    f = w.fields()
    e2e.release_cached_memory()
    assert e2e.cached_device_bytes() == 0
    outs = []
    for _ in range(3):
        with e2e.build_model(w.params, f) as m:
            csr = m.export_csr()
            pol, val, st = m.solve(thresh=1e-6)
            outs.append((csr, pol, val, st["sweeps"]))
        # the closed model's large blocks (entries, staging, values) stay mapped
        assert e2e.cached_device_bytes() > 0
    for csr, pol, val, sw in outs[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(csr, outs[0][0]))
        assert np.array_equal(pol, outs[0][1]) and sw == outs[0][3]
        assert np.array_equal(val.view(np.uint32), outs[0][2].view(np.uint32))
    e2e.release_cached_memory()
    assert e2e.cached_device_bytes() == 0
