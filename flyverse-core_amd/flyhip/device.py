"""Device choice and the sparse-matrix factory: the surface of flyverse/device.py:15-45 for ROCm.

PyTorch-ROCm reports HIP devices as device.type == "cuda", so `default_device` picks the GPU first exactly as the
reference does.  `sparse_matrix` returns the backend's own row-major `hip.CSR` (deterministic chunked summation
order) on a ROCm GPU; on the CPU it returns the torch CSR tensor the reference builds, for host-side callers only.
"""
from __future__ import annotations

import os

import numpy as np
import scipy.sparse as sp
import torch


def default_device() -> torch.device:
    """FLYVERSE_DEVICE if set, else the ROCm GPU, else CPU (flyverse/device.py:15-24)."""
    env = os.environ.get("FLYVERSE_DEVICE")
    if env:
        return torch.device(env)
    if torch.cuda.is_available():
        return torch.device("cuda")
    return torch.device("cpu")


def resolve(device=None) -> torch.device:
    if device is None:
        return default_device()
    return torch.device(device)


def sparse_matrix(D: sp.spmatrix, device, dtype=torch.float32):
    """scipy sparse -> a product-ready matrix on `device` (flyverse/device.py:31-45)."""
    device = torch.device(device)
    if device.type == "cuda":
        from .hip import CSR
        return CSR(D, device, dtype)
    C = sp.csr_matrix(D)
    return torch.sparse_csr_tensor(torch.from_numpy(C.indptr.astype(np.int64)),
                                   torch.from_numpy(C.indices.astype(np.int64)),
                                   torch.from_numpy(np.asarray(C.data)).to(dtype), size=C.shape)
