"""Host <-> device helpers for the GPU parity tests."""
from __future__ import annotations

import numpy as np
import torch


def to_prim(eb, state: np.ndarray, index_bits: int = 32):
    ts = [torch.from_numpy(np.ascontiguousarray(state[f])).cuda() for f in range(5)]
    return eb.EulerPrimArrays(*ts, index_bits=index_bits)


def to_cons(eb, state: np.ndarray, index_bits: int = 32):
    ts = [torch.from_numpy(np.ascontiguousarray(state[f])).cuda() for f in range(5)]
    return eb.EulerConsArrays(*ts, index_bits=index_bits)


def to_host(arrays) -> np.ndarray:
    torch.cuda.synchronize()
    return np.stack([t.cpu().numpy() for t in arrays.tensors()])


def max_rel(a: np.ndarray, b: np.ndarray) -> float:
    scale = np.maximum(np.abs(b), np.finfo(b.dtype).tiny)
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / scale))
