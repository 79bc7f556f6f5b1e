"""tq_grid_integrate_noalloc (block partials in a static device array) against
tq_grid_integrate with a caller workspace, through the C ABI."""
import ctypes
import warnings

import numpy as np
import pytest
import torch

from torchquad_amd import _lib, _ops
from torchquad_amd.integrands import GaussianPeak, Oscillatory

pytestmark = pytest.mark.gpu

CASES = [
    # (dtype, rule, dim, n, integrand) -- each pair: a larger call, then a smaller one
    (torch.float32, _lib.RULE_BOOLE, 5, 33, GaussianPeak([3.5, 2.1, 1.3, 4.2, 2.8], [0.1, -0.3, 0.2, 0.0, -0.4])),
    (torch.float32, _lib.RULE_SIMPSON, 3, 17, Oscillatory([1.2, 0.4, 1.9], 0.3)),
    (torch.float64, _lib.RULE_BOOLE, 4, 65, GaussianPeak([2.0, 3.0, 1.5, 2.5], [0.2, -0.1, 0.0, 0.3])),
    (torch.float64, _lib.RULE_TRAPEZOID, 2, 9, Oscillatory([0.7, 1.1], 0.1)),
]


def _run(L, entry, g, desc, dom, dev):
    out = torch.full((), float("nan"), dtype=dom.dtype, device=dev)
    out64 = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    if entry == "noalloc":
        rc = L.tq_grid_integrate_noalloc(ctypes.byref(g), ctypes.byref(desc), _lib.ptr(dom), _lib.ptr(out),
                                         _lib.ptr(out64), stream)
    else:
        nbytes = L.tq_grid_workspace_size(ctypes.byref(g))
        ws = torch.zeros(nbytes, dtype=torch.uint8, device=dev)  # the completion counter starts at zero
        rc = L.tq_grid_integrate(ctypes.byref(g), ctypes.byref(desc), _lib.ptr(dom), _lib.ptr(ws), nbytes,
                                 _lib.ptr(out), _lib.ptr(out64), stream)
    assert rc == 0, L.tq_last_error()
    torch.cuda.synchronize(dev)
    return out.item(), out64.item()


def test_noalloc_matches_workspace_entry_point():
    dev = torch.device("cuda", 0)
    L = _lib.load_library()
    for dtype, rule, dim, n, f in CASES:
        rng = np.random.default_rng(dim * 100 + n)
        lo = rng.uniform(-1.0, -0.2, dim)
        dom = torch.tensor(np.stack([lo, lo + rng.uniform(0.8, 1.6, dim)], 1), dtype=dtype, device=dev)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            g = _ops.grid_desc(rule, dim, n, dtype)
        desc = f.descriptor(dim)
        print(f"noalloc case dtype={dtype} rule={rule} dim={dim} n={n} family={desc.family}", flush=True)
        a = _run(L, "noalloc", g, desc, dom, dev)
        b = _run(L, "workspace", g, desc, dom, dev)
        assert a == b, (a, b)
        assert np.isfinite(a[1]) and a[1] != 0.0
        # the sum the result was formed from in FP64 agrees with the dtype output
        assert abs(a[0] - a[1]) <= (1e-6 if dtype == torch.float32 else 1e-15) * abs(a[1])


def test_noalloc_two_calls_second_smaller():
    dev = torch.device("cuda", 0)
    L = _lib.load_library()
    f_big = GaussianPeak([2.0, 3.0, 1.5, 2.5, 1.0], [0.2, -0.1, 0.0, 0.3, -0.2])
    f_small = GaussianPeak([4.0, 2.0], [0.1, -0.2])
    dom_big = torch.tensor([[-1.0, 1.0]] * 5, dtype=torch.float32, device=dev)
    dom_small = torch.tensor([[-0.5, 1.0], [0.0, 2.0]], dtype=torch.float32, device=dev)
    g_big = _ops.grid_desc(_lib.RULE_BOOLE, 5, 65, torch.float32)
    g_small = _ops.grid_desc(_lib.RULE_SIMPSON, 2, 21, torch.float32)
    first = _run(L, "noalloc", g_big, f_big.descriptor(5), dom_big, dev)
    second = _run(L, "noalloc", g_small, f_small.descriptor(2), dom_small, dev)
    assert first == _run(L, "workspace", g_big, f_big.descriptor(5), dom_big, dev)
    assert second == _run(L, "workspace", g_small, f_small.descriptor(2), dom_small, dev)
    # and the large one again after the small one left its partials behind
    assert first == _run(L, "noalloc", g_big, f_big.descriptor(5), dom_big, dev)
