"""Level passes whose items exceed the launch's LDS: not yet run on a GPU, so
kept under a name pytest does not collect.  Run it by naming the file:

    python -u -m pytest tests/wip_level_fallback.py -m gpu -x -v --timeout 120 \
        --timeout-method thread

At 800 bins every 32-row item of a three- or two-level pass stages more
than 64 KB of leaf rows, so the plan gives the level launch no LDS and every
item merges from global memory (clamped phases, the same FP32 sum tree); at
400 bins some items fit and some do not, so one launch takes both paths.
The detection tiles still fit.  The transform rows must equal the oracle's
bit for bit.
"""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows_bins", [(600, 800), (1500, 400)])
def test_level_items_from_global_memory_bit_exact(gpu, port, rows_bins):
    import gaffa_amd
    from oracle.oracle import TASK_DTYPE

    rows, bins = rows_bins
    print(f"ffa: transform rows={rows} bins={bins}, level items from global memory", flush=True)
    sys.stdout.flush()
    n = rows * bins + 3
    x = np.random.default_rng(rows + bins).normal(size=n).astype(np.float32)
    t = np.zeros(1, TASK_DTYPE)
    t[0] = (1.0, 1e-3, n, n, bins, rows, rows, bins * 1e-3, (bins + 1) * 1e-3)
    plan = gpu.FfaPlan(t, [1])
    layout = gpu._ffa_level_layout(plan)
    lib = ctypes.CDLL(gaffa_amd.HIP_LIBRARY)
    f = lib.gaffa_hip_ffa_level_lds_bytes
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint32] * 6
    needs = [f(rows, bins, a & 0xFFFF, a >> 16, first, count)
             for _task, a, first, count in layout["items"].tolist()]
    cap = 64 * 1024
    assert len(needs) > 0 and max(needs) > cap  # some items merge from global memory
    assert layout["level_lds_bytes"] == max([x for x in needs if x <= cap], default=0)
    got = gpu._ffa_transform_rows_gpu(x[None, :], plan, 0)[0]
    want = port.ffa_transform(x[: rows * bins], bins)
    np.testing.assert_array_equal(got, want)
