"""Device plumbing: torch owns HBM buffers and supplies the HIP stream; all
arithmetic happens in libcmm_s2_amd.so."""
import numpy as np
import torch

from .. import _native as N

F64 = torch.float64


def device():
    if not torch.cuda.is_available():
        raise N.CmmError("cmm_s2 needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_cur_dev = getattr(torch._C, "_cuda_getDevice", None)


def stream():
    """Handle of torch's current HIP stream on the current device (the
    capture stream during graph capture).  Called once per launch, so it
    takes torch's direct accessors when they exist (~10x cheaper than
    torch.cuda.current_stream())."""
    if _raw_stream is not None and _cur_dev is not None:
        return _raw_stream(_cur_dev())
    return torch.cuda.current_stream().cuda_stream


def is_numpy(x):
    return isinstance(x, np.ndarray)


def to_dev(x, dtype=F64):
    """Contiguous device tensor of dtype (copies host arrays once)."""
    dev = device()
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=dtype)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x), dtype=np.float64 if dtype == F64 else None)).to(dev, dtype)
    return t.contiguous()


def like(t, ref):
    """Return t as the caller's array type (numpy in -> numpy out)."""
    if is_numpy(ref):
        return t.detach().cpu().numpy()
    return t


def empty(*shape, dtype=F64):
    return torch.empty(*shape, dtype=dtype, device=device())


def zeros(*shape, dtype=F64):
    return torch.zeros(*shape, dtype=dtype, device=device())


def ptr(t):
    return None if t is None else t.data_ptr()


def point_strides(t, layout):
    """(stride between points, stride between components) in elements for a
    2-D tensor of points in layout 'N3' ((N,3)) or '3N' ((3,N))."""
    if layout == "N3":
        return t.stride(0), t.stride(1)
    return t.stride(1), t.stride(0)


_host_only = set()


def call_plugin(fn, *args):
    """fn(*args) for a caller-provided field (tensor arguments on the
    device): called on the device tensors first; a callable that fails there
    or returns no tensor (the reference's fields use np.* functions) is
    evaluated from then on with host NumPy copies of the tensor arguments
    (warned once), the result moved back to the device as float64."""
    dev = next(a.device for a in args if isinstance(a, torch.Tensor))
    if id(fn) not in _host_only:
        try:
            r = fn(*args)
        except Exception:
            r = None
        if isinstance(r, torch.Tensor):
            return r.to(device=dev, dtype=torch.float64)
        import warnings
        warnings.warn(f"{getattr(fn, '__name__', fn)!r} does not take device tensors: it is evaluated on the "
                      "host (NumPy) with a device round trip per call", RuntimeWarning, stacklevel=3)
        _host_only.add(id(fn))
    host = [a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a for a in args]
    r = np.asarray(fn(*host), dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(r)).to(dev)
