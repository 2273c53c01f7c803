"""Helpers shared by the GPU parity tests (torch is only device plumbing)."""
import numpy as np
import torch


def to_dev(x: np.ndarray) -> torch.Tensor:
    x = np.ascontiguousarray(x)
    if x.dtype == np.uint32:
        x = x.view(np.int32)
    t = torch.from_numpy(x.copy()).cuda()
    torch.cuda.synchronize()
    return t


def from_dev(t: torch.Tensor, dtype=np.uint32) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)


def empty_dev(shape, dtype=torch.int32) -> torch.Tensor:
    t = torch.zeros(shape, dtype=dtype, device="cuda")
    torch.cuda.synchronize()
    return t


def natural_of_position(N: int, N1: int, N2: int = 256) -> np.ndarray:
    """Natural coefficient index of each column-major plane position f = r*N1 + s."""
    f = np.arange(N)
    return (f // N1) + N2 * (f % N1)


def to_planes(nat: np.ndarray, N1: int) -> np.ndarray:
    """[..., N] natural order -> column-major plane positions."""
    N = nat.shape[-1]
    return nat[..., natural_of_position(N, N1)]


def from_planes(pos: np.ndarray, N1: int) -> np.ndarray:
    N = pos.shape[-1]
    out = np.empty_like(pos)
    out[..., natural_of_position(N, N1)] = pos
    return out


def sync():
    torch.cuda.synchronize()
