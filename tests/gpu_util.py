"""Helpers shared by the GPU tests: device copies and error norms."""
import numpy as np
import torch

import picasso_amd as pa

DEV = "cuda:0"


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def soa(a, dtype=torch.float64):
    """[P, q] numpy -> SoA device view: a [P, q] tensor with strides (1, P)."""
    t = torch.as_tensor(np.ascontiguousarray(np.asarray(a).T), dtype=dtype, device=DEV)
    return t.t()


def host(t):
    return t.detach().cpu().numpy()


def rel_inf(got, ref):
    """||got - ref||_inf / ||ref||_inf."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.abs(ref).max()
    return float(np.abs(got - ref).max() / (den if den > 0 else 1.0))


_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        _CTX = pa.Context(0)
    return _CTX
