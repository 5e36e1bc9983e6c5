"""kmsd_moments_noalloc (no device allocation; fixed per-device scratch)
against kmsd_moments through the C ABI: bitwise equal moments on two calls in
a row, the second with a smaller block, no bytes charged to the context, and
clear errors for a block beyond the scratch and for a nonidentity metric."""
import ctypes as C

import numpy as np
import pytest

from helpers import dense_model, hip_engine, host, oracle_engine, phases, rel_err, to_dev
from kubomsd_amd.synthetic import honeycomb_lattice

pytestmark = pytest.mark.gpu


def noalloc():
    from kubomsd_amd.backends import hip
    fn = hip.library().kmsd_moments_noalloc
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_double, C.c_double]
    fn.restype = C.c_int
    return fn


def call(ctx, eng, B, L, c, a):
    import torch
    from kubomsd_amd.backends.hip import _dtype_code
    mu = torch.full((L, B.shape[1]), np.nan, dtype=torch.float64, device=B.device)
    ctx.bind_stream()
    rc = noalloc()(eng.h, B.data_ptr(), mu.data_ptr(), L, B.shape[1], _dtype_code(B), c, a)
    return rc, mu


def test_noalloc_moments_equal_moments_twice(hip_ctx):
    m = honeycomb_lattice(40, 30, t=-2.7, W=1.0)
    eng = hip_engine(hip_ctx, m)
    oe = oracle_engine(m)
    for nb, L, dtype in ((16, 300, "c64"), (5, 77, "c128")):  # the second call is smaller
        B0 = phases(m.n, nb, 31 + nb)
        B = to_dev(hip_ctx, B0, dtype)
        want = host(eng.moments(B, L, 0.05, 8.8))
        used = hip_ctx.memory()["used_bytes"]
        rc, mu = call(hip_ctx, eng, B, L, 0.05, 8.8)
        assert rc == 0, hip_ctx.error()
        assert hip_ctx.memory()["used_bytes"] == used
        np.testing.assert_array_equal(host(mu), want)
        assert rel_err(host(mu), oe.moments(B0, L, 0.05, 8.8)) < (1e-11 if dtype == "c128" else 5e-7)


def test_noalloc_rejects_oversized_block_and_nonidentity_metric(hip_ctx):
    import torch
    from kubomsd_amd.backends.hip import KMSD_EINVAL
    m = honeycomb_lattice(40, 30, t=-2.7, W=1.0)
    eng = hip_engine(hip_ctx, m)
    # 2400 rows x 4096 complex128 columns = 157 MB per block: two do not fit 256 MiB
    B = torch.zeros((m.n, 4096), dtype=torch.complex128, device=hip_ctx.torch_device)
    rc, _ = call(hip_ctx, eng, B, 8, 0.0, 8.8)
    assert rc == KMSD_EINVAL and "fixed scratch" in hip_ctx.error()
    d = dense_model(9, 903)
    rc, _ = call(hip_ctx, hip_engine(hip_ctx, d), to_dev(hip_ctx, phases(9, 2, 1)), 8, 0.0, 5.0)
    assert rc == KMSD_EINVAL and "identity metric" in hip_ctx.error()
