"""MI355X-native FFT engine: Python mirror of wgpuFFT's plan surface.

The names, argument meaning and error behaviour follow the reference's public
API (src/config.rs, src/tuning.rs, src/plan.rs, src/error.rs); every call goes
through the C ABI of ``libmifft.so`` (include/mifft.h).  Buffers are device
pointers: torch CUDA/HIP tensors (their ``data_ptr()`` and byte size are used)
or ``(pointer, nbytes)`` pairs.  Streams are ``torch.cuda.Stream`` objects or
raw ``hipStream_t`` integers; the default is torch's current stream.

There is no CPU fallback: if ``libmifft.so`` is missing, importing works but
creating a plan raises ``LibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os
import enum
from typing import Iterable, Optional, Sequence

from . import _lib
from ._lib import LibraryMissing

__all__ = [
    "FftPrecision", "FftDirection", "Normalization", "FftTransformKind", "AxisKind", "C2cRoute",
    "FftTuning", "FftConfig", "FftPlan", "FftRecorder", "FftBatch", "FftError", "LibraryMissing",
    "create_plan", "create_r2c_plan", "create_c2r_plan", "fill_uniform_f32", "fill_normal_f32",
    "FftLogicalLayout", "FftPlanCreationError", "CpuFftPlan", "set_cpu_threads", "cpu_threads", "cpu_isa",
]


class FftPrecision(enum.IntEnum):
    """config.rs:5-17; complex element sizes 8 (F32), 16 (F64, Df64)."""
    F32 = 0
    F64 = 1
    Df64 = 2

    def scalar_size_bytes(self) -> int:
        return 4 if self == FftPrecision.F32 else 8

    def complex_size_bytes(self) -> int:
        return 2 * self.scalar_size_bytes()


class FftDirection(enum.IntEnum):
    Forward = 0
    Inverse = 1


class Normalization(enum.IntEnum):
    NONE = 0
    Forward = 1
    Inverse = 2
    Orthogonal = 3


class FftTransformKind(enum.IntEnum):
    C2c = 0
    R2c = 1
    C2r = 2


class AxisKind(enum.IntEnum):
    Mixed = 0
    Rader = 1
    Bluestein = 2


class C2cRoute(enum.IntEnum):
    DirectDft = 0
    MixedRadix = 1
    Rader = 2
    Bluestein = 3
    AxisSequence = 4


# mifft_status -> FftError variant name (src/error.rs:12-180).
_STATUS_NAMES = {
    1: "ZeroLength", 2: "ZeroBatch", 3: "LengthTooLarge", 4: "UnsupportedLength",
    5: "PrecisionUnsupported", 6: "InvalidTuning", 7: "EmptyAxes", 8: "InvalidAxis",
    9: "DuplicateAxis", 10: "InvalidRealTransformDirection", 11: "UnsupportedRealAxes",
    12: "BufferViewTooSmall", 13: "WorkspaceTooSmall", 14: "InPlaceUnsupported",
    15: "GpuPlanResourceAllocationFailed", 16: "HostBufferLength", 17: "NullArgument",
    18: "InvalidEnum", 19: "RankTooLarge", 20: "Hip", 21: "RecorderState", 22: "OverBudget",
    23: "Busy", 24: "BufferViewOffsetUnaligned", 25: "BufferLayoutZeroStride",
    26: "BufferLayoutBatchStrideTooSmall", 27: "BufferLayoutOutOfBounds", 28: "BufferLayoutTooLarge",
}


class FftError(Exception):
    """FftError: ``.kind`` is the reference variant name, ``.code`` the C status."""

    def __init__(self, code: int, detail: str = ""):
        self.code = int(code)
        self.kind = _STATUS_NAMES.get(self.code, f"Status{code}")
        self.detail = detail
        self.diagnostics = None  # FftExecutionError / FftPlanCreationError diagnostics, when recorded
        super().__init__(f"{self.kind}: {detail}" if detail else self.kind)


def _last_diagnostics():
    """The thread's last plan-creation / execution failure diagnostics, or None."""
    import json
    lib = _lib.load()
    n = ctypes.c_size_t()
    lib.mifft_last_error_diagnostics(None, 0, ctypes.byref(n))
    buf = ctypes.create_string_buffer(n.value)
    lib.mifft_last_error_diagnostics(buf, n.value, ctypes.byref(n))
    return json.loads(buf.value.decode())


def _check(status: int) -> None:
    if status != 0:
        lib = _lib.load()
        err = FftError(status, (lib.mifft_last_error_detail() or b"").decode())
        err.diagnostics = _last_diagnostics()
        raise err


class FftPlanCreationError(FftError):
    """FftPlanCreationError (error.rs:187-240): the error and its diagnostics."""

    def error(self) -> "FftError":
        return self

    def diagnostics_dict(self) -> dict:
        return self.diagnostics


class FftTuning:
    """FftTuning (tuning.rs:35-77): reference defaults, builder-style setters."""

    def __init__(self):
        self.workgroup_size = 64
        self.fused_workgroup_size = 256
        self.rader_max_prime = 8192
        self.direct_max_prime = 127
        self.force_rader_axes: list[int] = []
        self.force_bluestein_axes: list[int] = []
        self.fuse_long_axes = True
        self.fuse_small_volumes = True
        self.fused_min_convolution_length = 128

    @classmethod
    def default(cls) -> "FftTuning":
        return cls()

    def _with(self, **kw) -> "FftTuning":
        t = FftTuning()
        t.__dict__.update(self.__dict__)
        t.force_rader_axes = list(self.force_rader_axes)
        t.force_bluestein_axes = list(self.force_bluestein_axes)
        t.__dict__.update(kw)
        return t

    def with_workgroup_size(self, v: int): return self._with(workgroup_size=v)
    def with_fused_workgroup_size(self, v: int): return self._with(fused_workgroup_size=v)
    def with_rader_max_prime(self, v: int): return self._with(rader_max_prime=v)
    def with_direct_max_prime(self, v: int): return self._with(direct_max_prime=v)
    def with_force_rader_axes(self, axes: Iterable[int]): return self._with(force_rader_axes=list(axes))
    def with_force_bluestein_axes(self, axes: Iterable[int]): return self._with(force_bluestein_axes=list(axes))
    def with_fuse_long_axes(self, v: bool): return self._with(fuse_long_axes=bool(v))
    def with_fuse_small_volumes(self, v: bool): return self._with(fuse_small_volumes=bool(v))
    def with_fused_min_convolution_length(self, v: int): return self._with(fused_min_convolution_length=v)

    def _to_c(self) -> _lib.Tuning:
        t = _lib.Tuning()
        t.workgroup_size = self.workgroup_size
        t.fused_workgroup_size = self.fused_workgroup_size
        t.rader_max_prime = self.rader_max_prime
        t.direct_max_prime = self.direct_max_prime
        if len(self.force_rader_axes) > _lib.MAX_FORCED or len(self.force_bluestein_axes) > _lib.MAX_FORCED:
            raise FftError(6, "too many forced axes")
        for i, a in enumerate(self.force_rader_axes):
            t.force_rader_axes[i] = a
        t.n_force_rader_axes = len(self.force_rader_axes)
        for i, a in enumerate(self.force_bluestein_axes):
            t.force_bluestein_axes[i] = a
        t.n_force_bluestein_axes = len(self.force_bluestein_axes)
        t.fuse_long_axes = int(self.fuse_long_axes)
        t.fuse_small_volumes = int(self.fuse_small_volumes)
        t.fused_min_convolution_length = self.fused_min_convolution_length
        return t

    def __eq__(self, other):
        return isinstance(other, FftTuning) and self.__dict__ == other.__dict__


class FftConfig:
    """FftConfig (config.rs:75-287)."""

    def __init__(self, shape: Sequence[int]):
        self._shape = [int(s) for s in shape]
        self._axes = list(range(len(self._shape)))
        self._batch = 1
        self._direction = FftDirection.Forward
        self._normalization = Normalization.Inverse
        self._precision = FftPrecision.F32
        self._tuning = FftTuning()

    # constructors
    @classmethod
    def new(cls, n: int) -> "FftConfig":
        return cls([n])

    @classmethod
    def new_nd(cls, shape: Sequence[int]) -> "FftConfig":
        return cls(shape)

    @classmethod
    def inverse(cls, n: int) -> "FftConfig":
        return cls([n]).with_direction(FftDirection.Inverse)

    @classmethod
    def inverse_nd(cls, shape: Sequence[int]) -> "FftConfig":
        return cls(shape).with_direction(FftDirection.Inverse)

    def _copy(self) -> "FftConfig":
        c = FftConfig(self._shape)
        c._axes = list(self._axes)
        c._batch = self._batch
        c._direction = self._direction
        c._normalization = self._normalization
        c._precision = self._precision
        c._tuning = self._tuning
        return c

    def with_direction(self, d): c = self._copy(); c._direction = FftDirection(d); return c
    def with_normalization(self, n): c = self._copy(); c._normalization = Normalization(n); return c
    def with_axes(self, axes): c = self._copy(); c._axes = [int(a) for a in axes]; return c
    def with_batch(self, b): c = self._copy(); c._batch = int(b); return c
    def with_precision(self, p): c = self._copy(); c._precision = FftPrecision(p); return c
    def with_tuning(self, t): c = self._copy(); c._tuning = t; return c

    def len(self) -> int: return self._shape[0] if self._shape else 0
    def is_empty(self) -> bool: return not self._shape or 0 in self._shape
    def shape(self): return list(self._shape)
    def axes(self): return list(self._axes)
    def batch(self): return self._batch
    def direction(self): return self._direction
    def normalization(self): return self._normalization
    def precision(self): return self._precision
    def tuning(self): return self._tuning

    def _to_c(self) -> _lib.Config:
        c = _lib.Config()
        if len(self._shape) > _lib.MAX_RANK or len(self._axes) > _lib.MAX_RANK:
            raise FftError(19, "rank > 8")
        for i, s in enumerate(self._shape):
            if s < 0:
                raise FftError(1, "negative length")
            c.shape[i] = s
        c.rank = len(self._shape)
        for i, a in enumerate(self._axes):
            c.axes[i] = a
        c.naxes = len(self._axes)
        if self._batch < 0:
            raise FftError(2, "negative batch")
        c.batch = self._batch
        c.direction = int(self._direction)
        c.normalization = int(self._normalization)
        c.precision = int(self._precision)
        c.tuning = self._tuning._to_c()
        return c

    def validate(self) -> None:
        c = self._to_c()
        _check(_lib.load().mifft_config_validate(ctypes.byref(c)))

    def total_complex_len(self) -> int:
        c = self._to_c()
        v = ctypes.c_uint64()
        _check(_lib.load().mifft_config_total_len(ctypes.byref(c), ctypes.byref(v)))
        return v.value

    def logical_complex_len(self) -> int:
        p = 1
        for s in self._shape:
            p *= s
        return p

    def required_scalar_len(self) -> int:
        return 2 * self.total_complex_len()

    required_f32_len = required_scalar_len

    def required_buffer_size_bytes(self) -> int:
        return self.total_complex_len() * self._precision.complex_size_bytes()

    def scale_f64(self) -> float:
        c = self._to_c()
        v = ctypes.c_double()
        _check(_lib.load().mifft_config_scale(ctypes.byref(c), ctypes.byref(v)))
        return v.value

    def scale(self) -> float:
        import numpy as np
        return float(np.float32(self.scale_f64()))


def _buf(x, nbytes: Optional[int] = None):
    """(pointer, nbytes) of a torch tensor, an int pointer, or a (ptr, n) pair."""
    if isinstance(x, tuple):
        return int(x[0]), int(x[1])
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr()), int(x.numel() * x.element_size())
    if isinstance(x, int):
        if nbytes is None:
            raise FftError(17, "raw pointer without a size")
        return x, int(nbytes)
    raise TypeError(f"unsupported buffer type {type(x)!r}")


def _stream(s) -> int:
    if s is None:
        import torch
        return int(torch.cuda.current_stream().cuda_stream)
    if hasattr(s, "cuda_stream"):
        return int(s.cuda_stream)
    return int(s)


class FftLogicalLayout:
    """FftLogicalLayout (src/runtime/logical_io.rs:33-38): logical element i of
    batch b sits at element ``element_offset + b * batch_stride + i *
    element_stride`` of its buffer (elements of the endpoint's format);
    ``batch_stride=None`` is the per-batch span."""

    def __init__(self, element_offset: int = 0, element_stride: int = 1, batch_stride: Optional[int] = None):
        self.element_offset = int(element_offset)
        self.element_stride = int(element_stride)
        self.batch_stride = None if batch_stride is None else int(batch_stride)

    @classmethod
    def contiguous(cls) -> "FftLogicalLayout":
        return cls()

    def with_batch_stride(self, batch_stride: int) -> "FftLogicalLayout":
        return FftLogicalLayout(self.element_offset, self.element_stride, batch_stride)

    def _c(self):
        l = _lib.LogicalLayout()
        l.element_offset, l.element_stride = self.element_offset, self.element_stride
        l.batch_stride = self.batch_stride or 0
        return l

    def required_element_span(self, per_batch: int, batch: int) -> int:
        """Elements the layout spans (FftLogicalLayout::required_element_span)."""
        n = ctypes.c_uint64()
        c = self._c()
        _check(_lib.load().mifft_logical_layout_span(ctypes.byref(c), int(per_batch), int(batch), ctypes.byref(n)))
        return int(n.value)


def _layout_ptr(layout: Optional[FftLogicalLayout]):
    return None if layout is None else ctypes.byref(layout._c())


class FftPlan:
    """FftPlan (plan.rs:32-936) over HIP device buffers."""

    def __init__(self, config: FftConfig, kind: FftTransformKind):
        lib = _lib.load()
        self._config = config
        self._kind = FftTransformKind(kind)
        c = config._to_c()
        h = ctypes.c_void_p()
        _check(lib.mifft_plan_create(ctypes.byref(c), int(kind), ctypes.byref(h)))
        self._h = h
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mifft_plan_destroy(h)
            self._h = None

    # constructors (plan.rs:221,327,355)
    @classmethod
    def c2c(cls, config: FftConfig) -> "FftPlan":
        return cls(config, FftTransformKind.C2c)

    @classmethod
    def r2c(cls, config: FftConfig) -> "FftPlan":
        return cls(config, FftTransformKind.R2c)

    @classmethod
    def c2r(cls, config: FftConfig) -> "FftPlan":
        return cls(config, FftTransformKind.C2r)

    @classmethod
    def _with_diagnostics(cls, config: FftConfig, kind: FftTransformKind) -> "FftPlan":
        import json
        lib = _lib.load()
        c = config._to_c()
        h = ctypes.c_void_p()
        n = ctypes.c_size_t()
        buf = ctypes.create_string_buffer(1 << 16)
        st = lib.mifft_plan_create_with_diagnostics(ctypes.byref(c), int(kind), ctypes.byref(h), buf, 1 << 16,
                                                    ctypes.byref(n))
        if st != 0:
            err = FftPlanCreationError(st, (lib.mifft_last_error_detail() or b"").decode())
            err.diagnostics = json.loads(buf.value.decode())
            raise err
        plan = cls.__new__(cls)
        plan._config, plan._kind, plan._h, plan._lib = config, FftTransformKind(kind), h, lib
        return plan

    # FftPlan::c2c_with_diagnostics (plan.rs:227-235) and its real counterparts:
    # a failure raises FftPlanCreationError carrying the diagnostics
    @classmethod
    def c2c_with_diagnostics(cls, config: FftConfig) -> "FftPlan":
        return cls._with_diagnostics(config, FftTransformKind.C2c)

    @classmethod
    def r2c_with_diagnostics(cls, config: FftConfig) -> "FftPlan":
        return cls._with_diagnostics(config, FftTransformKind.R2c)

    @classmethod
    def c2r_with_diagnostics(cls, config: FftConfig) -> "FftPlan":
        return cls._with_diagnostics(config, FftTransformKind.C2r)

    def diagnostics_for_workspace(self, input, output, workspace) -> dict:
        """FftPlan::diagnostics_for_workspace (plan.rs:1023-1055)."""
        import json
        ip, ib = _buf(input)
        op, ob = _buf(output)
        wp, wb = _buf(workspace)
        n = ctypes.c_size_t()
        buf = ctypes.create_string_buffer(1 << 16)
        _check(self._lib.mifft_plan_diagnostics_for_workspace(self._h, ip, ib, op, ob, wp, wb, buf, 1 << 16,
                                                              ctypes.byref(n)))
        return json.loads(buf.value.decode())

    def diagnostics_for_views(self, input, input_layout: Optional["FftLogicalLayout"], output,
                              output_layout: Optional["FftLogicalLayout"], workspace=None) -> dict:
        """FftPlan::diagnostics_for_views / _for_logical_views(_with_workspace)
        (plan.rs:1057-1300); `workspace` a buffer, a byte count, or None for
        the plan's own."""
        import json
        ip, ib = _buf(input)
        op, ob = _buf(output)
        wsz = None
        if workspace is not None:
            wsz = ctypes.c_uint64(int(workspace) if isinstance(workspace, int) else _buf(workspace)[1])
        n = ctypes.c_size_t()
        buf = ctypes.create_string_buffer(1 << 16)
        _check(self._lib.mifft_plan_diagnostics_for_views(
            self._h, ip, ib, _layout_ptr(input_layout), op, ob, _layout_ptr(output_layout),
            ctypes.byref(wsz) if wsz is not None else None, buf, 1 << 16, ctypes.byref(n)))
        return json.loads(buf.value.decode())

    # queries (plan.rs:390-492)
    def kind(self) -> FftTransformKind: return self._kind
    def config(self) -> FftConfig: return self._config

    def axis_factors(self) -> list[list[int]]:
        out = []
        buf = (ctypes.c_uint64 * 64)()
        for i in range(len(self._config.axes())):
            n = self._lib.mifft_plan_axis_factors(self._h, i, buf, 64)
            out.append([int(buf[k]) for k in range(max(n, 0))])
        return out

    def factors(self) -> list[int]:
        return self.axis_factors()[0]

    def axis_kinds(self) -> list[AxisKind]:
        buf = (ctypes.c_int32 * 8)()
        n = self._lib.mifft_plan_axis_kinds(self._h, buf, 8)
        return [AxisKind(buf[i]) for i in range(n)]

    def route(self) -> C2cRoute: return C2cRoute(self._lib.mifft_plan_route(self._h))
    def workspace_size_bytes(self) -> int: return int(self._lib.mifft_plan_workspace_bytes(self._h))
    def required_input_buffer_size_bytes(self) -> int: return int(self._lib.mifft_plan_required_input_bytes(self._h))
    def required_output_buffer_size_bytes(self) -> int: return int(self._lib.mifft_plan_required_output_bytes(self._h))
    def required_buffer_size_bytes(self) -> int: return int(self._lib.mifft_plan_required_bytes(self._h))

    def packed_shape(self) -> Optional[list[int]]:
        buf = (ctypes.c_uint64 * 8)()
        n = self._lib.mifft_plan_packed_shape(self._h, buf, 8)
        return [int(buf[i]) for i in range(n)] if n else None

    def supports_in_place(self) -> bool: return bool(self._lib.mifft_plan_supports_in_place(self._h))
    def launch_count(self) -> int: return int(self._lib.mifft_plan_launch_count(self._h))

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(8192)
        _check(self._lib.mifft_plan_describe(self._h, buf, 8192))
        return buf.value.decode()

    def launch_describe(self, index: int) -> str:
        buf = ctypes.create_string_buffer(1024)
        _check(self._lib.mifft_plan_launch_describe(self._h, int(index), buf, 1024))
        return buf.value.decode()

    def tables(self) -> list:
        """(description dict, numpy uint8 array) of every plan-owned device
        table (mifft_plan_table)."""
        import numpy as np
        out = []
        d = ctypes.create_string_buffer(512)
        for i in range(int(self._lib.mifft_plan_table_count(self._h))):
            nb = ctypes.c_uint64()
            _check(self._lib.mifft_plan_table(self._h, i, d, 512, None, 0, ctypes.byref(nb)))
            host = np.zeros(nb.value, np.uint8)
            _check(self._lib.mifft_plan_table(self._h, i, d, 512, host.ctypes.data_as(ctypes.c_void_p), nb.value,
                                              ctypes.byref(nb)))
            words = d.value.decode().split()
            desc = {"kind": words[0]}
            for w in words[1:]:
                k, v = w.split("=", 1)
                desc[k] = v
            out.append((desc, host))
        return out

    def launch_tuning(self, index: int) -> tuple:
        """(tuning-table key, registered variant name) of launch `index`;
        either is "" where it does not apply."""
        k = ctypes.create_string_buffer(256)
        v = ctypes.create_string_buffer(256)
        _check(self._lib.mifft_plan_launch_tuning(self._h, int(index), k, 256, v, 256))
        return k.value.decode(), v.value.decode()

    def diagnostics(self) -> dict:
        """FftPlan::diagnostics (plan.rs:938-1004): route, stages, blockers,
        buffer requirements, active tuning and device limits."""
        import json
        n = ctypes.c_size_t()
        _check(self._lib.mifft_plan_diagnostics(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(self._lib.mifft_plan_diagnostics(self._h, buf, n.value, ctypes.byref(n)))
        return json.loads(buf.value.decode())

    def execute_launch(self, index: int, input, output, workspace=None, stream=None) -> None:
        """Enqueue only launch `index` of one execution (per-kernel timing)."""
        ip, ib = _buf(input)
        op, ob = _buf(output)
        wp, wb = _buf(workspace) if workspace is not None else (None, 0)
        _check(self._lib.mifft_execute_launch(self._h, _stream(stream), int(index), ip, ib, op, ob, wp, wb))

    # execution (plan.rs:494-836)
    def execute_checked(self, input, output, stream=None) -> None:
        ip, ib = _buf(input)
        op, ob = _buf(output)
        _check(self._lib.mifft_execute(self._h, _stream(stream), ip, ib, op, ob))

    def execute(self, input, output, stream=None) -> None:
        self.execute_checked(input, output, stream)

    def execute_with_workspace(self, input, output, workspace, stream=None) -> None:
        ip, ib = _buf(input)
        op, ob = _buf(output)
        wp, wb = _buf(workspace)
        _check(self._lib.mifft_execute_with_workspace(self._h, _stream(stream), ip, ib, op, ob, wp, wb))

    def execute_in_place(self, buffer, stream=None) -> None:
        p, b = _buf(buffer)
        _check(self._lib.mifft_execute_in_place(self._h, _stream(stream), p, b))

    # logical views (plan.rs:734-806, 917-936)
    def logical_workspace_size_bytes(self, input_layout: Optional[FftLogicalLayout] = None,
                                     output_layout: Optional[FftLogicalLayout] = None) -> int:
        n = ctypes.c_uint64()
        _check(self._lib.mifft_plan_logical_workspace_bytes(self._h, _layout_ptr(input_layout),
                                                            _layout_ptr(output_layout), ctypes.byref(n)))
        return int(n.value)

    def execute_logical_views(self, input, input_layout: Optional[FftLogicalLayout], output,
                              output_layout: Optional[FftLogicalLayout], stream=None) -> None:
        ip, ib = _buf(input)
        op, ob = _buf(output)
        _check(self._lib.mifft_execute_logical(self._h, _stream(stream), ip, ib, _layout_ptr(input_layout),
                                               op, ob, _layout_ptr(output_layout)))

    def execute_logical_views_with_workspace(self, input, input_layout: Optional[FftLogicalLayout], output,
                                             output_layout: Optional[FftLogicalLayout], workspace,
                                             stream=None) -> None:
        ip, ib = _buf(input)
        op, ob = _buf(output)
        wp, wb = _buf(workspace)
        _check(self._lib.mifft_execute_logical_with_workspace(self._h, _stream(stream), ip, ib,
                                                              _layout_ptr(input_layout), op, ob,
                                                              _layout_ptr(output_layout), wp, wb))

    def record_logical_views(self, recorder: "FftRecorder", input, input_layout: Optional[FftLogicalLayout],
                             output, output_layout: Optional[FftLogicalLayout]) -> None:
        ip, ib = _buf(input)
        op, ob = _buf(output)
        _check(self._lib.mifft_record_logical(recorder._h, self._h, ip, ib, _layout_ptr(input_layout), op, ob,
                                              _layout_ptr(output_layout)))

    def record(self, recorder: "FftRecorder", input, output) -> None:
        recorder._record(self, input, output)

    def record_in_place(self, recorder: "FftRecorder", buffer) -> None:
        recorder._record(self, buffer, buffer)


class CpuFftPlan:
    """CpuFftPlan (src/cpu.rs:52-361): the same configuration and layouts on
    host memory (numpy arrays of scalar words: float32 for F32 and Df64 plans,
    float64 for F64 plans)."""

    def __init__(self, config: FftConfig, kind: FftTransformKind):
        lib = _lib.load()
        self._config = config
        self._kind = FftTransformKind(kind)
        c = config._to_c()
        h = ctypes.c_void_p()
        _check(lib.mifft_cpu_plan_create(ctypes.byref(c), int(kind), ctypes.byref(h)))
        self._h = h
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mifft_cpu_plan_destroy(h)
            self._h = None

    @classmethod
    def c2c(cls, config: FftConfig) -> "CpuFftPlan":
        return cls(config, FftTransformKind.C2c)

    @classmethod
    def r2c(cls, config: FftConfig) -> "CpuFftPlan":
        return cls(config, FftTransformKind.R2c)

    @classmethod
    def c2r(cls, config: FftConfig) -> "CpuFftPlan":
        return cls(config, FftTransformKind.C2r)

    def kind(self) -> FftTransformKind: return self._kind
    def config(self) -> FftConfig: return self._config

    def packed_shape(self) -> Optional[list[int]]:
        buf = (ctypes.c_uint64 * _lib.MAX_RANK)()
        n = self._lib.mifft_cpu_plan_packed_shape(self._h, buf, _lib.MAX_RANK)
        return [int(buf[i]) for i in range(n)] if n else None

    def required_input_len(self) -> int: return int(self._lib.mifft_cpu_plan_required_input_len(self._h))
    def required_output_len(self) -> int: return int(self._lib.mifft_cpu_plan_required_output_len(self._h))

    @staticmethod
    def _words(a, dtype, writable=False):
        import numpy as np
        if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags["C_CONTIGUOUS"] or \
                (writable and not a.flags["WRITEABLE"]):
            raise TypeError(f"expected a contiguous{' writable' if writable else ''} {np.dtype(dtype).name} array")
        return a.ctypes.data_as(ctypes.c_void_p), a.size

    def execute(self, input, output) -> None:
        """F32 and Df64 plans on float32 words (cpu.rs:245-268)."""
        import numpy as np
        ip, il = self._words(input, np.float32)
        op, ol = self._words(output, np.float32, True)
        _check(self._lib.mifft_cpu_execute(self._h, ip, il, op, ol))

    def execute_f64(self, input, output) -> None:
        """F64 plans on float64 words (cpu.rs:270-283)."""
        import numpy as np
        ip, il = self._words(input, np.float64)
        op, ol = self._words(output, np.float64, True)
        _check(self._lib.mifft_cpu_execute_f64(self._h, ip, il, op, ol))

    def execute_in_place(self, data) -> None:
        import numpy as np
        p, n = self._words(data, np.float32, True)
        _check(self._lib.mifft_cpu_execute_in_place(self._h, p, n))

    def execute_in_place_f64(self, data) -> None:
        import numpy as np
        p, n = self._words(data, np.float64, True)
        _check(self._lib.mifft_cpu_execute_in_place_f64(self._h, p, n))


def set_cpu_threads(n: int) -> None:
    """Host threads CpuFftPlan executions may use (0: every hardware thread)."""
    _lib.load().mifft_cpu_set_max_threads(int(n))


def cpu_threads() -> int:
    return int(_lib.load().mifft_cpu_max_threads())


def cpu_isa() -> str:
    return _lib.load().mifft_cpu_isa().decode()


class FftRecorder:
    """FftRecorder (plan.rs:38-95): records executions onto one stream, or
    captures them into one hipGraph (``capture_graph=True``) replayed by
    :meth:`launch`.  Use as a context manager; leaving it ends recording."""

    def __init__(self, stream=None, capture_graph: bool = False):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _check(self._lib.mifft_recorder_begin(_stream(stream), int(capture_graph), ctypes.byref(h)))
        self._h = h
        self._ended = False

    def compute_pass(self) -> int:
        """The stream executions are recorded on (caller work goes here)."""
        return int(self._lib.mifft_recorder_stream(self._h) or 0)

    def _record(self, plan: FftPlan, input, output) -> None:
        ip, ib = _buf(input)
        op, ob = _buf(output)
        _check(self._lib.mifft_record(self._h, plan._h, ip, ib, op, ob))

    def end(self) -> None:
        if not self._ended:
            self._ended = True
            _check(self._lib.mifft_recorder_end(self._h))

    def launch(self, stream=None) -> None:
        _check(self._lib.mifft_recorder_launch(self._h, _stream(stream)))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.end()
        return False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mifft_recorder_destroy(h)
            self._h = None


class FftBatch:
    """Many 1D f32 C2C transforms of uneven lengths in one persistent launch.

    ``items``: sequence of dicts with keys ``n``, ``batch``, ``input``,
    ``output`` (device buffers, may be the same) and optional ``direction``,
    ``normalization`` (default Inverse) and ``algorithm`` (0 route choice,
    1 force Rader, 2 force Bluestein).  A transform's result does not depend
    on the other items (bit for bit).
    """

    def __init__(self, items, tuning: Optional[FftTuning] = None):
        self._lib = _lib.load()
        arr = (_lib.BatchItem * len(items))()
        self._keep = []
        for k, it in enumerate(items):
            ip, _ = _buf(it["input"])
            op, _ = _buf(it["output"])
            arr[k].n = int(it["n"])
            arr[k].batch = int(it["batch"])
            arr[k].direction = int(it.get("direction", FftDirection.Forward))
            arr[k].normalization = int(it.get("normalization", Normalization.Inverse))
            arr[k].algorithm = int(it.get("algorithm", 0))
            arr[k].in_ptr = ip
            arr[k].out_ptr = op
            self._keep.append((it["input"], it["output"]))
        t = (tuning or FftTuning())._to_c()
        h = ctypes.c_void_p()
        _check(self._lib.mifft_batch_create(arr, len(items), ctypes.byref(t), ctypes.byref(h)))
        self._h = h

    def execute(self, stream=None) -> None:
        _check(self._lib.mifft_batch_execute(self._h, _stream(stream)))

    def capture(self, stream=None) -> None:
        _check(self._lib.mifft_batch_capture(self._h, _stream(stream)))

    def launch(self, stream=None) -> None:
        _check(self._lib.mifft_batch_launch(self._h, _stream(stream)))

    def info(self) -> dict:
        t, p, l, w = (ctypes.c_uint64() for _ in range(4))
        g = ctypes.c_uint32()
        _check(self._lib.mifft_batch_info(self._h, ctypes.byref(t), ctypes.byref(p), ctypes.byref(l),
                                          ctypes.byref(w), ctypes.byref(g)))
        return {"tiles": t.value, "phases": p.value, "lds_bytes": l.value, "workspace_bytes": w.value,
                "grid": g.value}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mifft_batch_destroy(h)
            self._h = None


def create_plan(config: FftConfig) -> FftPlan:
    return FftPlan.c2c(config)


def create_r2c_plan(config: FftConfig) -> FftPlan:
    return FftPlan.r2c(config)


def create_c2r_plan(config: FftConfig) -> FftPlan:
    return FftPlan.c2r(config)


# ------------------------------------------------------------------ autotuner
def autotune(config: FftConfig, kind: FftTransformKind = FftTransformKind.C2c, reps: int = 20,
             stream=None) -> tuple:
    """Times every candidate kernel of each line pass of the plan of (config,
    kind) and records the fastest in the process-wide tuning table, which
    plans created afterwards use.  Returns (passes tuned, report text)."""
    lib = _lib.load()
    c = config._to_c()
    n = ctypes.c_int32()
    buf = ctypes.create_string_buffer(1 << 16)
    _check(lib.mifft_autotune(ctypes.byref(c), int(kind), int(reps), _stream(stream), ctypes.byref(n), buf,
                              len(buf)))
    return n.value, buf.value.decode()


def autotune_sequence(jobs, reps: int = 20, stream=None) -> tuple:
    """Tunes the kernel shapes of several plans executed back to back, as a
    caller runs them (a forward+inverse pair in place; R2C then C2R chained).
    `jobs` is a list of (FftConfig, FftTransformKind).  Returns (shapes tuned,
    report text)."""
    lib = _lib.load()
    cs = (_lib.Config * len(jobs))(*[cfg._to_c() for cfg, _ in jobs])
    ks = (ctypes.c_int32 * len(jobs))(*[int(k) for _, k in jobs])
    n = ctypes.c_int32()
    buf = ctypes.create_string_buffer(1 << 16)
    _check(lib.mifft_autotune_sequence(cs, ks, len(jobs), int(reps), _stream(stream), ctypes.byref(n), buf,
                                       len(buf)))
    return n.value, buf.value.decode()


def twiddle_table(tensor, n: int, count: int, precision: "FftPrecision" = None, stream=None) -> None:
    """W_n^k for k < count into a device buffer (mifft_twiddle_table)."""
    p, b = _buf(tensor)
    prec = FftPrecision.F32 if precision is None else FftPrecision(precision)
    if b < count * prec.complex_size_bytes():
        raise ValueError("buffer too small for the table")
    _check(_lib.load().mifft_twiddle_table(p, int(n), int(count), int(prec), _stream(stream)))


def registry_names() -> list:
    """Names of every compiled register-kernel variant (tuning-table choices)."""
    lib = _lib.load()
    out = []
    buf = ctypes.create_string_buffer(256)
    for i in range(int(lib.mifft_registry_count())):
        _check(lib.mifft_registry_name(i, buf, 256))
        out.append(buf.value.decode())
    return out


def save_tuning(path: str) -> None:
    _check(_lib.load().mifft_tuning_table_save(os.fsencode(path)))


def load_tuning(path: str) -> None:
    _check(_lib.load().mifft_tuning_table_load(os.fsencode(path)))


def parse_tuning(text: str) -> int:
    n = ctypes.c_int32()
    _check(_lib.load().mifft_tuning_table_parse(text.encode(), ctypes.byref(n)))
    return n.value


def clear_tuning() -> None:
    _lib.load().mifft_tuning_table_clear()


def tuning_size() -> int:
    return int(_lib.load().mifft_tuning_table_size())


def fill_uniform_f32(tensor, seed: int, stream=None) -> None:
    """SplitMix64 uniform[-1,1) f32 fill on the device (element k = draw k)."""
    p, b = _buf(tensor)
    _check(_lib.load().mifft_fill_uniform_f32(p, b // 4, seed, _stream(stream)))


def fill_normal_f32(tensor, seed: int, stream=None) -> None:
    """SplitMix64 + Box-Muller N(0,1) f32 fill on the device."""
    p, b = _buf(tensor)
    _check(_lib.load().mifft_fill_normal_f32(p, b // 4, seed, _stream(stream)))
