"""Device plumbing: pick the ROCm device, allocate torch tensors.

PyTorch only provides device memory and the stream; all physics runs in
libssim.  Every entry point that computes calls `require_gpu` first, which
raises when no GPU is visible or when libssim is not built — there is no
host fallback.
"""
from __future__ import annotations

import atexit
import gc
import sys
import weakref

import numpy as np

from . import _lib

# Objects whose close() releases GPU resources that must go while the HIP
# runtime is still up (captured graphs, kernel-written pinned buffers).
_closables: "weakref.WeakSet" = weakref.WeakSet()


def on_shutdown(obj) -> None:
    """Have obj.close() run by shutdown() (at the latest at interpreter
    exit, before modules and the HIP runtime are torn down)."""
    _closables.add(obj)


def shutdown() -> None:
    """Release the library's GPU state in a defined order: wait for the
    device, close every registered object (graph executables, pinned host
    stages, workspaces), collect what they referenced, wait again and hand
    the allocator's cached blocks back.  Safe to call more than once; also
    registered with atexit."""
    torch = sys.modules.get("torch")
    live = torch is not None and torch.cuda.is_initialized()
    if live:
        torch.cuda.synchronize()
    for obj in list(_closables):
        obj.close()
    _closables.clear()
    gc.collect()
    if live:
        torch.cuda.synchronize()
        torch.cuda.empty_cache()


atexit.register(shutdown)


def require_gpu(device: int = 0):
    """Return torch.device(f'cuda:{device}') or raise if it cannot run."""
    import torch
    _lib.load()
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm GPU visible: libssim needs an MI355X")
    if not 0 <= device < torch.cuda.device_count():
        raise ValueError(f"device {device} not present "
                         f"({torch.cuda.device_count()} visible)")
    return torch.device("cuda", device)


def to_dev(a, dev, dtype=None):
    """Host array -> contiguous device tensor (float64 by default)."""
    import torch
    dtype = dtype or torch.float64
    arr = np.ascontiguousarray(a)
    return torch.as_tensor(arr).to(device=dev, dtype=dtype).contiguous()


def empty(n: int, dev, dtype=None):
    import torch
    return torch.empty(max(int(n), 1), dtype=dtype or torch.float64, device=dev)


def zeros(n: int, dev, dtype=None):
    import torch
    return torch.zeros(max(int(n), 1), dtype=dtype or torch.float64, device=dev)


def workspace(nbytes: int, dev):
    """Byte buffer for a libssim workspace (None when no scratch is needed)."""
    import torch
    if nbytes <= 0:
        return None
    return torch.empty(int(nbytes), dtype=torch.uint8, device=dev)


def stream(dev) -> int:
    return _lib.stream_handle(dev)
