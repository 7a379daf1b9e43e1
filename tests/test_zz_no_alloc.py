"""The allocation-free backprojection entry point (vcbct_reconstruct_image_noalloc) against the allocating one
(vcbct_reconstruct_image), both called through the C ABI: two calls in a row on the same GPU, the second with a
smaller geometry (fewer projections, a smaller panel and volume), each equal to the allocating entry point's
result; sizes beyond the table's maximum are rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def _case(torch, R, Z, n, H, W, seed):
    rng = np.random.default_rng(seed)
    ang = np.deg2rad(-180.0 + np.arange(n) * (360.0 / n) + 1e-4)
    xs = np.full(n, 0.37)
    # a smooth field plus noise, as a filtered projection set looks
    yy, xx = np.mgrid[0:H, 0:W]
    base = np.cos(xx / W * 7.0) * np.exp(-((yy - H / 2) / H) ** 2)
    sino = (base[None] + 0.1 * rng.standard_normal((n, H, W))).astype(np.float32)
    return ang, xs, torch.from_numpy(sino).cuda()


def _both(torch, R, Z, zstart, step, scale, n, H, W, ang, xs, sino):
    from virtualcbct_amd._lib import P_D, check, lib
    L = lib()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    ref = torch.zeros((Z, R, R), dtype=torch.float32, device="cuda")
    a = np.ascontiguousarray(ang)
    x = np.ascontiguousarray(xs)
    check(L.vcbct_reconstruct_image(R, Z, zstart, step, scale, C.c_void_p(ref.data_ptr()), H, W, n,
                                    C.c_void_p(sino.data_ptr()), 1000, a.ctypes.data_as(P_D), x.ctypes.data_as(P_D),
                                    0.032509752, st))
    got = torch.zeros_like(ref)
    da = torch.from_numpy(a).cuda()
    dx = torch.from_numpy(x).cuda()
    check(L.vcbct_reconstruct_image_noalloc(R, Z, zstart, step, scale, C.c_void_p(got.data_ptr()), H, W, n,
                                            C.c_void_p(sino.data_ptr()), 1000, C.c_void_p(da.data_ptr()),
                                            C.c_void_p(dx.data_ptr()), 0.032509752, st))
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_backprojection_equals_allocating_entry_point(torch):
    for (R, Z, n, H, W, seed) in ((270, 16, 360, 400, 400, 1), (64, 4, 36, 128, 96, 2)):
        print(f"noalloc backprojection: R={R} Z={Z} n={n} panel {H}x{W}", flush=True)
        ang, xs, sino = _case(torch, R, Z, n, H, W, seed)
        ref, got = _both(torch, R, Z, -1, 0.1, 0.1, n, H, W, ang, xs, sino)
        level = float(np.abs(ref).max())
        diff = float(np.abs(got - ref).max())
        print(f"  max |noalloc - allocating| = {diff:.3e} of level {level:.3e}; "
              f"{int((got != ref).sum())} of {ref.size} voxels differ", flush=True)
        assert level > 0
        assert diff <= 1e-6 * level


def test_noalloc_rejects_oversized_tables(torch):
    from virtualcbct_amd._lib import VCBCT_EINVAL, lib
    L = lib()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    rec = torch.zeros(16, dtype=torch.float32, device="cuda")
    p = C.c_void_p(buf.data_ptr())
    rc = L.vcbct_reconstruct_image_noalloc(4, 1, 0, 0.1, 0.1, C.c_void_p(rec.data_ptr()), 4, 4, 20000,
                                           C.c_void_p(rec.data_ptr()), 1000, p, p, 0.0325, None)
    assert rc == VCBCT_EINVAL
    assert b"maximum" in L.vcbct_last_error()
