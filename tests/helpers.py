"""Host <-> device helpers for the tests (column-major layout)."""
import numpy as np


def to_dev(a, dtype=None):
    """numpy (rows, cols) -> column-major CUDA tensor of the same shape."""
    import torch

    a = np.asarray(a if dtype is None else np.asarray(a, dtype=dtype))
    if a.ndim == 1:
        return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda").t()


def to_host(t):
    return t.detach().cpu().numpy()
