"""FP64 oracle for the FFT hot path -- TEST INFRASTRUCTURE ONLY.

This package is the reference the tests compare against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import it.  The product
(``wgpufft_amd`` and ``libmifft.so``) never imports, links or calls it.

It wraps ``oracle/liboracle.so`` (built from ``oracle/oracle.c`` by
``oracle/Makefile``), a C restatement of wgpuFFT's host reference transforms:

* ``c2c_nd``   -- ``reference_c2c_nd_f64``            (src/math/mod.rs:214-268)
* ``fft1d``    -- ``fft_f64`` / ``bluestein_f64``     (src/math/mod.rs:274-405)
* ``r2c_nd``   -- ``reference_r2c_packed_interleaved`` (src/math/mod.rs:81-112)
* ``c2r_nd``   -- ``reference_c2r_from_packed_interleaved`` (src/math/mod.rs:114-153)
* integer helpers of the route choice (src/runtime/mod.rs:29-54,
  src/runtime/axis_policy.rs:83-177, src/runtime/rader_axis.rs:1270-1304,
  src/runtime/bluestein_axis.rs:858-867,1035-1041)
* df64 split/join (src/cpu.rs:489-517)

Arrays are numpy; complex data is complex128 in C order with axis 0 of the
wgpuFFT shape as the *fastest* axis, i.e. a wgpuFFT shape ``[N0, N1]`` with
batch ``B`` is the flat buffer of ``B*N1*N0`` complex values.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

NORM_NONE, NORM_FORWARD, NORM_INVERSE, NORM_ORTHO = 0, 1, 2, 3


def build(force: bool = False) -> str:
    """Compile oracle/oracle.c with gcc (a few seconds)."""
    src = os.path.join(_HERE, "oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "liboracle.so"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    c_dp = ctypes.POINTER(ctypes.c_double)
    c_fp = ctypes.POINTER(ctypes.c_float)
    c_szp = ctypes.POINTER(ctypes.c_size_t)
    c_ip = ctypes.POINTER(ctypes.c_int)
    sz = ctypes.c_size_t
    L.oracle_fft1d.argtypes = [c_dp, c_dp, sz, ctypes.c_int]
    L.oracle_dft_naive.argtypes = [c_dp, c_dp, sz, ctypes.c_int]
    L.oracle_c2c_nd.argtypes = [c_dp, c_dp, c_szp, ctypes.c_int, c_ip, ctypes.c_int, sz,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.oracle_c2c_nd.restype = ctypes.c_int
    L.oracle_r2c_nd.argtypes = [c_dp, c_dp, c_szp, ctypes.c_int, sz, ctypes.c_int, ctypes.c_int]
    L.oracle_c2r_nd.argtypes = [c_dp, c_dp, c_szp, ctypes.c_int, sz, ctypes.c_int, ctypes.c_int]
    L.oracle_scale.argtypes = [c_szp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.oracle_scale.restype = ctypes.c_double
    L.oracle_factor_supported.argtypes = [sz, c_szp, ctypes.c_int]
    L.oracle_factor_supported.restype = ctypes.c_int
    L.oracle_is_prime.argtypes = [sz]
    L.oracle_is_prime.restype = ctypes.c_int
    L.oracle_mod_pow.argtypes = [sz, sz, sz]
    L.oracle_mod_pow.restype = sz
    L.oracle_primitive_root.argtypes = [sz]
    L.oracle_primitive_root.restype = sz
    L.oracle_next_smooth.argtypes = [sz]
    L.oracle_next_smooth.restype = sz
    L.oracle_rader_conv_len.argtypes = [sz]
    L.oracle_rader_conv_len.restype = sz
    L.oracle_rader_perm.argtypes = [sz, ctypes.POINTER(ctypes.c_uint32)]
    L.oracle_rader_perm.restype = ctypes.c_int
    L.oracle_bluestein_conv_len.argtypes = [sz]
    L.oracle_bluestein_conv_len.restype = sz
    L.oracle_bluestein_phase_index.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.oracle_bluestein_phase_index.restype = ctypes.c_uint64
    L.oracle_df64_split.argtypes = [c_dp, c_fp, sz]
    L.oracle_df64_join.argtypes = [c_fp, c_dp, sz]
    L.oracle_splitmix64_uniform_f32.argtypes = [ctypes.c_uint64, c_fp, sz]
    L.oracle_splitmix64_normal_f32.argtypes = [ctypes.c_uint64, c_fp, sz]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _shape_arr(shape):
    return (ctypes.c_size_t * len(shape))(*[int(s) for s in shape])


def fft1d(x, inverse: bool = False) -> np.ndarray:
    """Unnormalised 1D DFT (fft_f64, src/math/mod.rs:274)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().oracle_fft1d(_dp(x), _dp(out), x.size, 1 if inverse else -1)
    return out


def dft_naive(x, inverse: bool = False) -> np.ndarray:
    """Brute-force long-double DFT (pinning aid, O(n^2))."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().oracle_dft_naive(_dp(x), _dp(out), x.size, 1 if inverse else -1)
    return out


def c2c_nd(x, shape, axes=None, batch=1, inverse=False, normalization=NORM_INVERSE,
           nthreads=1) -> np.ndarray:
    """reference_c2c_nd_f64 on a flat complex buffer (axis 0 fastest)."""
    shape = [int(s) for s in shape]
    axes = list(range(len(shape))) if axes is None else [int(a) for a in axes]
    x = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
    assert x.size == int(np.prod(shape)) * batch
    out = np.empty_like(x)
    ax = (ctypes.c_int * len(axes))(*axes)
    rc = lib().oracle_c2c_nd(_dp(x), _dp(out), _shape_arr(shape), len(shape), ax, len(axes),
                             batch, 1 if inverse else 0, normalization, nthreads)
    assert rc == 0
    return out


def packed_shape(shape):
    s = [int(v) for v in shape]
    s[0] = s[0] // 2 + 1
    return s


def r2c_nd(real, shape, batch=1, normalization=NORM_INVERSE, nthreads=1) -> np.ndarray:
    """reference_r2c_packed_interleaved in f64; returns the packed complex buffer."""
    real = np.ascontiguousarray(real, dtype=np.float64).reshape(-1)
    assert real.size == int(np.prod(shape)) * batch
    out = np.empty(int(np.prod(packed_shape(shape))) * batch, dtype=np.complex128)
    lib().oracle_r2c_nd(_dp(real), _dp(out), _shape_arr(shape), len(shape), batch, normalization,
                        nthreads)
    return out


def c2r_nd(packed, shape, batch=1, normalization=NORM_INVERSE, nthreads=1) -> np.ndarray:
    """reference_c2r_from_packed_interleaved in f64; returns the real buffer."""
    packed = np.ascontiguousarray(packed, dtype=np.complex128).reshape(-1)
    assert packed.size == int(np.prod(packed_shape(shape))) * batch
    out = np.empty(int(np.prod(shape)) * batch, dtype=np.float64)
    lib().oracle_c2r_nd(_dp(packed), _dp(out), _shape_arr(shape), len(shape), batch, normalization,
                        nthreads)
    return out


def scale(shape, inverse, normalization) -> float:
    return lib().oracle_scale(_shape_arr(shape), len(shape), 1 if inverse else 0, normalization)


def factor_supported(n):
    buf = (ctypes.c_size_t * 64)()
    c = lib().oracle_factor_supported(n, buf, 64)
    return None if c < 0 else [int(buf[i]) for i in range(c)]


def is_prime(n):
    return bool(lib().oracle_is_prime(n))


def primitive_root(p):
    return int(lib().oracle_primitive_root(p)) or None


def next_smooth(n):
    return int(lib().oracle_next_smooth(n))


def rader_conv_len(n):
    return int(lib().oracle_rader_conv_len(n))


def rader_perm(n):
    out = (ctypes.c_uint32 * max(n - 1, 1))()
    rc = lib().oracle_rader_perm(n, out)
    return None if rc else [int(out[i]) for i in range(n - 1)]


def bluestein_conv_len(n):
    return int(lib().oracle_bluestein_conv_len(n))


def bluestein_phase_index(n, i):
    return int(lib().oracle_bluestein_phase_index(n, i))


def df64_split(v) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    out = np.empty(2 * v.size, dtype=np.float32)
    lib().oracle_df64_split(_dp(v), _fp(out), v.size)
    return out


def df64_join(hilo) -> np.ndarray:
    hilo = np.ascontiguousarray(hilo, dtype=np.float32).reshape(-1)
    out = np.empty(hilo.size // 2, dtype=np.float64)
    lib().oracle_df64_join(_fp(hilo), _dp(out), out.size)
    return out


def uniform_f32(seed, n) -> np.ndarray:
    out = np.empty(n, dtype=np.float32)
    lib().oracle_splitmix64_uniform_f32(seed, _fp(out), n)
    return out


def normal_f32(seed, n) -> np.ndarray:
    out = np.empty(n, dtype=np.float32)
    lib().oracle_splitmix64_normal_f32(seed, _fp(out), n)
    return out


def rel_l2(y, y_ref) -> float:
    """||y - y_ref||_2 / ||y_ref||_2 in f64 (rms_relative, tests/gpu_accuracy.rs:153-178)."""
    y = np.asarray(y, dtype=np.complex128).reshape(-1)
    y_ref = np.asarray(y_ref, dtype=np.complex128).reshape(-1)
    den = np.linalg.norm(y_ref)
    return float(np.linalg.norm(y - y_ref) / (den if den > 0 else 1.0))


def max_rel(y, y_ref) -> float:
    """max |y - y_ref| / max |y_ref| (the max-rel of tests/gpu_accuracy.rs)."""
    y = np.asarray(y, dtype=np.complex128).reshape(-1)
    y_ref = np.asarray(y_ref, dtype=np.complex128).reshape(-1)
    den = np.max(np.abs(y_ref)) if y_ref.size else 1.0
    return float(np.max(np.abs(y - y_ref)) / (den if den > 0 else 1.0)) if y.size else 0.0
