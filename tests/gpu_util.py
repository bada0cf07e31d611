"""Helpers shared by the GPU tests (device buffers are torch tensors)."""
import contextlib
import sys

import numpy as np
import torch


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def dev(x, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def empty(nbytes):
    return torch.zeros(max(int(nbytes), 16), dtype=torch.uint8, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@contextlib.contextmanager
def jacobi_method(runtime, method):
    """Run the block with this eigensolver (orders up to 256), then restore
    the library's default."""
    from msq import capi
    runtime.set_jacobi_method(method)
    try:
        yield runtime
    finally:
        runtime.set_jacobi_method(capi.JACOBI_DEFAULT)
