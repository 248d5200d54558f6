"""Tensor interop: host arrays or torch tensors in, device float tensors for the library."""
import numpy as np
import torch

DEVICE = "cuda"


def as_device(a, dtype=torch.float64):
    """Contiguous device tensor of `dtype` holding `a` (numpy array or torch tensor)."""
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=DEVICE, dtype=dtype).contiguous()
    np_dtype = {torch.float64: np.float64, torch.float32: np.float32, torch.uint8: np.uint8,
                torch.int64: np.int64}[dtype]
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np_dtype)).to(DEVICE)


def is_torch(a):
    return isinstance(a, torch.Tensor)


def output_like(t, ref):
    """Returns `t` as a numpy array when `ref` is not a torch tensor."""
    if is_torch(ref):
        return t
    return t.cpu().numpy()


def empty(n, dtype=torch.float64):
    return torch.empty(n, dtype=dtype, device=DEVICE)
