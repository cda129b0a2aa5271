"""Helpers for the GPU tests: device buffers through torch, calls through the C ABI."""
import ctypes as C

import numpy as np
import torch

import fastandfourier_amd as fa


def L():
    return fa.load()


def dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def empty(shape, dtype=torch.float32):
    return torch.empty(shape, dtype=dtype, device="cuda")


def host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


def create(kind: str, n: int, **kw):
    lib = L()
    cfg = fa.config_init(n, **kw)
    fn = {"fft": lib.faf_create_fft, "rfft": lib.faf_create_rfft, "stft": lib.faf_create_stft}[kind]
    t = fn(C.byref(cfg))
    if not t:
        raise fa.FafError(lib.faf_get_error().decode())
    return t


def destroy(t):
    L().faf_destroy_transform(t)


def c2c(t, x: np.ndarray, inter=False) -> np.ndarray:
    """Batched C2C over the rows of complex x (device split planes)."""
    lib = L()
    x = np.atleast_2d(x).astype(np.complex64)
    cnt, n = x.shape
    torch.cuda.synchronize()
    if inter:
        xi = dev(x.view(np.float32))
        yo = empty((cnt, 2 * n))
        bi = fa.buffer_interleaved(xi.data_ptr(), n)
        bo = fa.buffer_interleaved(yo.data_ptr(), n)
        fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, 0, 0), "c2c")
        return host(yo).view(np.complex64)
    xr, xi = dev(x.real.copy()), dev(x.imag.copy())
    yr, yi = empty((cnt, n)), empty((cnt, n))
    bi = fa.buffer_split(xr.data_ptr(), xi.data_ptr(), n)
    bo = fa.buffer_split(yr.data_ptr(), yi.data_ptr(), n)
    fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, 0, 0), "c2c")
    return host(yr) + 1j * host(yi)


def r2c(t, x: np.ndarray) -> np.ndarray:
    lib = L()
    x = np.atleast_2d(x).astype(np.float32)
    cnt, n = x.shape
    nb = n // 2 + 1
    dx = dev(x)
    re, im = empty((cnt, nb)), empty((cnt, nb))
    torch.cuda.synchronize()
    bi = fa.buffer_real(dx.data_ptr(), n)
    bo = fa.buffer_hermitian(re.data_ptr(), im.data_ptr(), nb)
    fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, 0, 0), "rfft")
    return host(re) + 1j * host(im)


def c2r(t, X: np.ndarray, n: int) -> np.ndarray:
    lib = L()
    X = np.atleast_2d(X).astype(np.complex64)
    cnt, nb = X.shape
    xr, xi = dev(X.real.copy()), dev(X.imag.copy())
    y = empty((cnt, n))
    torch.cuda.synchronize()
    bi = fa.buffer_hermitian(xr.data_ptr(), xi.data_ptr(), nb)
    bo = fa.buffer_real(y.data_ptr(), n)
    fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, 0, 0), "irfft")
    return host(y)


def real_to_real(t, x: np.ndarray) -> np.ndarray:
    """Pipelines: REAL in, REAL out, batched over rows."""
    lib = L()
    x = np.atleast_2d(x).astype(np.float32)
    cnt, n = x.shape
    dx = dev(x)
    y = empty((cnt, n))
    torch.cuda.synchronize()
    bi = fa.buffer_real(dx.data_ptr(), n)
    bo = fa.buffer_real(y.data_ptr(), n)
    fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, 0, 0), "pipeline")
    return host(y)
