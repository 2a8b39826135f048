"""closure_grad_noalloc (the entry point that allocates nothing) against
closure_grad (the caller's partial-gradient buffer), through the C ABI.

Both run the same kernels with the same number of workgroups, and the
partial sums are reduced in a fixed order, so the gradients and the loss must
agree bit for bit.  Two calls in a row, the second with fewer samples and
fewer workgroups: the library's array keeps the first call's rows, which the
second must not read.  The argument check beyond the array's size runs
without a GPU.
"""

import ctypes as C

import numpy as np
import pytest

from conftest import load_sub


def _inputs(rng, n, wdim):
    X = rng.standard_normal((n, 19)).astype(np.float32)
    T = (0.2 * rng.standard_normal((n, 4))).astype(np.float32)
    W = rng.uniform(0.5, 1.5, (n,) if wdim == 1 else (n, 4)).astype(np.float32)
    return X, T, W


@pytest.mark.gpu
@pytest.mark.parametrize("wdim", [1, 4])
def test_noalloc_gradient_equals_closure_grad(pkg, gpu, wdim):
    import torch
    lib = load_sub("_lib")
    cl = load_sub("closure")
    net = cl.ClosureNet(cl.init_params(5, out_scale=1.0), device=gpu)
    st = C.c_void_p(torch.cuda.current_stream(gpu).cuda_stream)
    P = lambda t: C.c_void_p(t.data_ptr())
    rng = np.random.default_rng(11)
    # first a large call, then a smaller one (fewer samples and workgroups)
    for n, blocks in ((5000, 48), (700, 5)):
        X, T, W = (torch.from_numpy(a).to(gpu) for a in _inputs(rng, n, wdim))
        wsum = C.c_float(float(W.double().sum().item()))
        partial = torch.full((lib.load().closure_partial_bytes(blocks) // 4,), float("nan"),
                             dtype=torch.float32, device=gpu)
        g_ref = torch.full((cl.NP + 1,), float("nan"), dtype=torch.float32, device=gpu)
        g_new = torch.full((cl.NP + 1,), float("nan"), dtype=torch.float32, device=gpu)
        lib.call("closure_grad", P(net.packed), P(net.params), n, P(X), P(T), P(W), wdim, wsum,
                 P(partial), blocks, P(g_ref), st)
        lib.call("closure_grad_noalloc", P(net.packed), P(net.params), n, P(X), P(T), P(W), wdim,
                 wsum, blocks, P(g_new), st)
        torch.cuda.synchronize(gpu)
        a, b = g_ref.cpu().numpy(), g_new.cpu().numpy()
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_noalloc_rejects_more_workgroups_than_its_array_holds():
    lib = load_sub("_lib")
    one = C.c_void_p(1)
    too_many = lib.CLOSURE_GRAD_NOALLOC_MAX_BLOCKS + 1
    with pytest.raises(lib.ClbmError, match="CLOSURE_GRAD_NOALLOC_MAX_BLOCKS"):
        lib.call("closure_grad_noalloc", one, one, 100, one, one, one, 1, C.c_float(1.0),
                 too_many, one, None)
    with pytest.raises(lib.ClbmError, match="NULL"):
        lib.call("closure_grad_noalloc", one, one, 100, None, one, one, 1, C.c_float(1.0), 4, one,
                 None)
