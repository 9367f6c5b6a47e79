"""Helpers shared by the GPU parity tests (host <-> device copies of field lists)."""
import numpy as np
import torch


def to_dev(arrs, device="cuda:0"):
    return [torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in arrs]


def to_host(ts):
    torch.cuda.synchronize()
    return [t.detach().cpu().numpy() for t in ts]
