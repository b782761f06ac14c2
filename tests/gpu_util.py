"""Device helpers for the GPU tests (torch tensors as plain device buffers)."""
import sys

import numpy as np


def torch():
    import torch as t
    return t


def dev(a, dtype):
    t = torch()
    return t.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype).contiguous()


def host(x):
    return x.detach().to("cpu").double().numpy()


def sync():
    torch().cuda.synchronize()


def stream():
    return torch().cuda.current_stream().cuda_stream


def say(*args):
    """Print a progress line naming the case before a GPU step, flushed."""
    print(*args, flush=True)
    sys.stdout.flush()
