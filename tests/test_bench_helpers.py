"""Host-side helpers of bench.py: the HBM traffic of the headline hop is read
from the committed rocprofv3 summary for exactly the launch shapes profiled
there, and the multi-RHS batches respect the kernels' 4 GB offset window."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402


def test_profiled_traffic_per_shape():
    a = bench.profiled_traffic((1, 1, 4, 1, 2))
    b = bench.profiled_traffic((1, 1, 4, 1, 0))
    # both above the 1.611 GB of compulsory bytes at 32^3 x 64, FP64, 18-real links
    compulsory = 1536 * 32**3 * 64 // 2
    assert a is not None and b is not None
    assert compulsory < b <= a < 1.2 * compulsory
    assert bench.profiled_traffic((1, 2, 4, 1, 0)) is None


def test_c4_batches_fit_the_slice_window_and_memory():
    vh_big, vh_small = 32**3 * 64 // 2, 16**3 * 32 // 2
    free = 280 << 30
    assert bench.c4_batch(vh_big, 12, 64, free) == 12
    assert bench.c4_batch(vh_big, 48, 64, free) == 48  # a 9.7 GB half field, addressed per t-slice
    assert bench.c4_batch(vh_big, 48, 64, 100 << 30) == 16  # memory caps it
    assert bench.c4_batch(vh_small, 48, 32, free) == 48
    for vh, n, lt in [(vh_big, 12, 64), (vh_big, 48, 64), (vh_small, 48, 32), (vh_big, 7, 64)]:
        nb = bench.c4_batch(vh, n, lt, free)
        assert n % nb == 0 and vh * 12 * 16 * nb // lt < 1 << 32 and 20 * vh * 12 * 16 * nb <= 0.85 * free
