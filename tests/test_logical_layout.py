"""Logical I/O layouts (FftLogicalLayout, src/runtime/logical_io.rs:33-178) on
the host: the element span a layout needs and its validation errors.  The span
is pinned by its closed form offset + (batch-1)*batch_stride + (n-1)*stride + 1
(logical_io.rs:123-145) and by brute-force enumeration of every element index."""
import pytest

import wgpufft_amd as W
from wgpufft_amd import FftLogicalLayout as L


@pytest.mark.parametrize("off,stride,bs,n,batch", [(0, 1, None, 10, 1), (0, 1, None, 10, 4), (7, 1, 12, 10, 3),
                                                   (3, 2, None, 5, 2), (5, 3, 40, 9, 4), (0, 4, 4 * 8, 8, 1),
                                                   (11, 1, 10, 10, 2)])
def test_span_matches_enumeration(off, stride, bs, n, batch):
    span = L(off, stride, bs).required_element_span(n, batch)
    bstride = bs if bs is not None else (n - 1) * stride + 1
    last = max(off + b * bstride + i * stride for b in range(batch) for i in range(n))
    assert span == last + 1
    assert span == off + (batch - 1) * bstride + (n - 1) * stride + 1


def test_layout_errors():
    with pytest.raises(W.FftError) as e:
        L(0, 0).required_element_span(4, 1)
    assert e.value.kind == "BufferLayoutZeroStride"
    with pytest.raises(W.FftError) as e:
        L(0, 2, 6).required_element_span(4, 2)  # per-batch span is 7
    assert e.value.kind == "BufferLayoutBatchStrideTooSmall"
    assert L(0, 2, 7).required_element_span(4, 2) == 14
    with pytest.raises(W.FftError) as e:
        L(1 << 62, 1 << 40).required_element_span(1 << 30, 2)
    assert e.value.kind == "BufferLayoutTooLarge"


def test_contiguous_defaults():
    c = L.contiguous()
    assert (c.element_offset, c.element_stride, c.batch_stride) == (0, 1, None)
    assert c.with_batch_stride(9).batch_stride == 9
    assert c.required_element_span(8, 3) == 24
