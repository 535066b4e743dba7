"""The device LBVH (Morton codes, radix sort, Karras hierarchy, hierarchical
refit) checked on the host after every build and refit: each leaf once,
parent and child links consistent, node boxes exactly the union of their
children's."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 4099, 100_000, 784_385])
def test_lbvh_build_and_refit(sc, n):
    ms = np.zeros(2)
    print(f"lbvh n={n}", flush=True)
    bad = sc.lib().scLbvhSelfCheck(n, 7 + n, 3, ms.ctypes.data_as(ctypes.c_void_p))
    print(f"  build {ms[0]:.3f} ms, refit {ms[1]:.3f} ms", flush=True)
    assert bad == 0
