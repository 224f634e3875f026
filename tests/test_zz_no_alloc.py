"""nw_weighted_sum_total (partials in the library's own device array, nothing allocated) against
nw_weighted_sum (partials in a caller's buffer, summed in order on the host): two calls in a row,
the second on a smaller input, and the size limit."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-no-alloc] {msg}", flush=True)
    sys.stdout.flush()


def test_weighted_sum_total_without_workspace(torch_gpu):
    from neuwon_amd import _lib
    torch = torch_gpu
    lib = _lib.load()
    P = _lib.ptr
    rng = np.random.default_rng(12)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    total = torch.zeros(1, dtype=torch.float64, device="cuda")
    for n, parts in ((1_000_003, 4096), (12_345, 300)):
        w = torch.as_tensor(rng.uniform(0.1, 2.0, n), device="cuda")
        x = torch.as_tensor(rng.uniform(0.0, 1.0, n).astype(np.float32), device="cuda")
        y = torch.as_tensor(rng.uniform(0.0, 100.0, n).astype(np.float32), device="cuda")
        z = torch.as_tensor(rng.normal(0.0, 1.0, n).astype(np.float32), device="cuda")
        partials = torch.zeros(parts, dtype=torch.float64, device="cuda")
        _lib.check(lib.nw_weighted_sum(n, P(w), P(x), P(y), P(z), 0.05, P(partials), parts, s))
        _lib.check(lib.nw_weighted_sum_total(n, P(w), P(x), P(y), P(z), 0.05, parts, P(total), s))
        torch.cuda.synchronize()
        want = 0.0
        for p in partials.cpu().numpy().tolist():   # in order, as the device sums them
            want += p
        got = float(total.item())
        exact = float(np.sum(w.cpu().numpy() * (x.cpu().numpy().astype(np.float64)
                                                 + y.cpu().numpy().astype(np.float64)
                                                 + 0.05 * z.cpu().numpy().astype(np.float64))))
        _say(f"n {n}, {parts} partials: total {got!r}, host sum of the partials {want!r}, "
             f"numpy FP64 {exact!r}")
        assert got == want
        assert abs(got - exact) <= 1e-12 * abs(exact)
    # without the optional arrays (y, z NULL)
    _lib.check(lib.nw_weighted_sum(n, P(w), P(x), None, None, 0.0, P(partials), parts, s))
    _lib.check(lib.nw_weighted_sum_total(n, P(w), P(x), None, None, 0.0, parts, P(total), s))
    torch.cuda.synchronize()
    want = 0.0
    for p in partials.cpu().numpy().tolist():
        want += p
    assert float(total.item()) == want
    # more partials than the array holds, or none: refused before any launch
    for bad in (_lib_max() + 1, 0):
        rc = lib.nw_weighted_sum_total(n, P(w), P(x), None, None, 0.0, bad, P(total), s)
        assert rc != 0 and b"num_partials" in lib.nw_last_error()


def _lib_max():
    return 4096   # NW_SUM_SCRATCH_PARTIALS
