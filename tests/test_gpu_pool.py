"""Device blocks kept mapped after a model closes (devpool.cpp) serve the next models:
builds and solves in reused memory give the same bits as in fresh memory, and the idle
blocks go back to the runtime on request."""
import numpy as np
import pytest

import e2e_gpu_dp_amd as e2e
from e2e_gpu_dp_amd.configs import workload
from e2e_gpu_dp_amd.fields import make_fields

pytestmark = pytest.mark.gpu


def _run(p, f, variant=0):
    with e2e.build_model(p, f) as m:
        csr = m.export_csr()
        pol, val, st = m.solve(thresh=1e-6, kernel_variant=variant)
    return csr, pol, val, st["sweeps"]


def _same(a, b):
    assert all(np.array_equal(x, y) for x, y in zip(a[0], b[0]))
    assert np.array_equal(a[1], b[1]) and a[3] == b[3]
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))


@pytest.mark.parametrize("variant", [0, 3])
def test_reused_device_memory_gives_the_same_models_and_solves(gpu, variant):
    """Two models of the same shape from different fields, built and solved in turn: the
    blocks each one reuses hold the other model's bytes (entries, staging, values), so a
    kernel that read memory it had not written would change the results."""
    w = workload("C1")
    p = w.params
    fa = w.fields()
    fb = make_fields(p.gsize, p.nt, p.nmodes, p.nrzns, w.seed + 17, w.sigmas)
    e2e.release_cached_memory()
    assert e2e.cached_device_bytes() == 0
    a0 = _run(p, fa, variant)  # fresh memory
    b0 = _run(p, fb, variant)
    # the two models differ (so reused memory holds different bytes)
    assert not np.array_equal(a0[2].view(np.uint32), b0[2].view(np.uint32))
    assert e2e.cached_device_bytes() > 0  # the closed models' large blocks stay mapped
    for _ in range(2):
        _same(_run(p, fa, variant), a0)
        _same(_run(p, fb, variant), b0)
    e2e.release_cached_memory()
    assert e2e.cached_device_bytes() == 0
