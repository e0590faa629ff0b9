"""mh_assemble_noalloc (fixed device scratch, no allocation in the call)
against mh_assemble, through the C ABI: two calls in a row, the second with
a smaller mesh and fewer wavenumbers, so stale scratch contents would show."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(mh):
    return mh._lib.lib().mh_assemble_noalloc


def _run(mh, torch, fn, body, ks, form):
    dm = mh.DeviceMesh(body.mesh)
    n = dm.nfaces
    nk = len(ks)
    k = np.ascontiguousarray(np.asarray(ks, dtype=np.float64))
    S = torch.zeros((nk, n, n), dtype=torch.complex128, device="cuda:0")
    D = torch.zeros_like(S) if form & 1 else None
    K = torch.zeros_like(S) if form & 2 else None
    rc = fn(mh._lib.context(0), C.byref(dm.struct), 1 | 2 | 8, form, nk, k.ctypes.data, S.data_ptr(),
            D.data_ptr() if D is not None else None, K.data_ptr() if K is not None else None, n, n * n)
    mh._lib.check(rc, fn.__name__)
    torch.cuda.synchronize()
    return [X.cpu().numpy() if X is not None else None for X in (S, D, K)]


@pytest.mark.parametrize("form", [1, 2, 3])
def test_noalloc_matches_assemble_twice(mh, gpu, form):
    torch = gpu
    na = _noalloc(mh)
    ref = mh._lib.lib().mh_assemble
    cases = [(mh.hemisphere(1.0, ntheta=24, nphi=20), [0.2, 0.7, 1.3, 2.9, 4.0]),
             (mh.hemisphere(1.3, ntheta=12, nphi=10), [0.5, 1.1])]
    for body, ks in cases:  # the second call is smaller than the first
        print(f"noalloc form={form} n={body.mesh.nfaces} nk={len(ks)}", flush=True)
        got = _run(mh, torch, na, body, ks, form)
        exp = _run(mh, torch, ref, body, ks, form)
        for g, e in zip(got, exp):
            if e is None:
                continue
            err = np.max(np.abs(g - e)) / np.max(np.abs(e))
            assert err < 1e-12, (form, body.mesh.nfaces, err)


def test_noalloc_rejects_sizes_beyond_the_scratch(mh, gpu):
    torch = gpu
    na = _noalloc(mh)
    body = mh.hemisphere(1.0, ntheta=8, nphi=8)
    with pytest.raises(ValueError, match="wavenumbers exceed"):
        _run(mh, torch, na, body, np.linspace(0.1, 3.0, 300), 1)
    big = mh.hemisphere(1.0, ntheta=132, nphi=64)  # 4224 panels
    dm = mh.DeviceMesh(big.mesh)
    k = np.array([1.0])
    rc = na(mh._lib.context(0), C.byref(dm.struct), 1 | 2 | 8, 1, 1, k.ctypes.data, 1, 1, None,
            dm.nfaces, dm.nfaces * dm.nfaces)
    assert rc == 1 and b"exceed the fixed scratch limit" in mh._lib.lib().mh_last_error()
