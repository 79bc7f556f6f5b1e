"""Gauss-Legendre nodes and weights from the device generator (tq_gauss_legendre)."""
import torch

from ._lib import NativeUnavailableError, call, ptr, stream_ptr


def gauss_legendre(n, device=None):
    """(nodes, weights) float64 on [-1, 1], ascending, computed on the GPU."""
    if not torch.cuda.is_available():
        raise NativeUnavailableError("Gauss-Legendre nodes are generated on a ROCm GPU only")
    device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    x = torch.empty(int(n), dtype=torch.float64, device=device)
    w = torch.empty(int(n), dtype=torch.float64, device=device)
    call("tq_gauss_legendre", int(n), ptr(x), ptr(w), stream_ptr(device))
    return x, w


def fp32_pairs(x, w):
    """Interleaved FP32 (node, weight) pairs of a float64 table: [x0, w0, x1, w1, ...].

    FP32 batches read one 8-byte pair per point instead of two 8-byte
    float64 loads; each value is the float64 one rounded to nearest."""
    return torch.stack((x.float(), w.float()), dim=1).reshape(-1)
