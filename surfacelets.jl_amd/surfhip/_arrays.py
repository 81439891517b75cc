"""Column-major device arrays (torch tensors) for the C ABI.

The reference's arrays are Julia column-major: a volume of size (n1, n2, n3)
keeps axis 1 contiguous.  Here a volume is a torch tensor of shape
(n1, n2, n3) whose strides are (1, n1, n1 n2); torch tensors of that layout
are views of a contiguous (n3, n2, n1) buffer, so no data moves.
"""
import numpy as np
import torch

from ._lib import C64, F32, ArgumentError


def colmajor_strides(shape):
    s, out = 1, []
    for n in shape:
        out.append(s)
        s *= n
    return tuple(out)


def is_colmajor(t):
    return tuple(t.stride()) == colmajor_strides(t.shape) or t.numel() <= 1


def empty_colmajor(shape, dtype, device):
    """An uninitialised column-major tensor of the given shape."""
    base = torch.empty(tuple(reversed(shape)), dtype=dtype, device=device)
    return base.permute(*reversed(range(len(shape))))


def as_colmajor(t):
    """`t` itself if column-major, else a column-major copy."""
    if is_colmajor(t):
        return t
    out = empty_colmajor(tuple(t.shape), t.dtype, t.device)
    out.copy_(t)
    return out


def from_numpy(a, device="cuda"):
    """Upload a numpy array (any layout) as a column-major device tensor."""
    a = np.asarray(a)
    base = torch.from_numpy(np.ascontiguousarray(np.transpose(a)))
    return base.to(device).permute(*reversed(range(a.ndim)))


def to_numpy(t):
    """Download a device tensor (any layout) to a numpy array of its shape."""
    return t.detach().cpu().numpy()


def data_dtype(t):
    """The library dtype code of a tensor, after the reference's rule that an
    integer element type means its float (here: Float32)."""
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.complex64:
        return C64
    raise ArgumentError(
        f"libsurfhip computes in Float32 and ComplexF32; got element type {t.dtype}")


def check_device(t, what="array"):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise ArgumentError(f"the {what} must be a device (HIP) tensor")


def stream_handle():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(v):
    import ctypes
    return ctypes.c_void_p(int(v))


def ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())
