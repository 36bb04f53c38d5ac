"""vj_integral_image_noalloc (the allocation-free single-image integral)
against the entry point it is a variant of, vj_engine_integral_batch with one
image, on two calls in a row into the same caller buffers, the second with a
smaller image, and against the oracle; oversized images are refused."""
import numpy as np
import pytest

from cs677s26_amd import synth
from cs677s26_amd._native import VJ_INVALID_ARG, check, lib
from cs677s26_amd.device import DeviceArray, Stream
from cs677s26_amd.engine import FaceVisionEngine
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def padded(img, stride):
    h, w = img.shape
    buf = np.full((h, stride), 0xA5, np.uint8)  # bytes past w must not be read
    buf[:, :w] = img
    return buf


def via_engine(img, stride):
    h, w = img.shape
    eng = FaceVisionEngine((24, 24), 1, w, h, False, 1000)
    d = DeviceArray.from_host(padded(img, stride))
    eng.computeIntegralBatchTransposed(d, w, h, stride, 1, False)
    s, q, _ = eng.integral_to_host()
    return s.reshape(h + 1, w + 1), q.reshape(h + 1, w + 1)


def test_noalloc_integral_equals_engine_on_two_calls():
    big = synth.detect_image(1920, 1080, 40, 24, 200, seed=71)
    small = synth.detect_image(641, 479, 10, 24, 80, seed=72)
    cases = [(big, 1984), (small, 700)]
    # caller buffers sized for the first call and reused by the second, filled
    # with a marker so that every output element must be written
    cap = (1080 + 1) * (1920 + 1)
    dS, dQ = DeviceArray(cap, np.uint32), DeviceArray(cap, np.uint64)
    st = Stream()
    for img, stride in cases:
        h, w = img.shape
        say(f"noalloc integral {w}x{h} stride {stride}")
        n = (h + 1) * (w + 1)
        check(lib().vj_memset(dS.ptr, 0xFF, dS.nbytes, st.handle), "vj_memset")
        check(lib().vj_memset(dQ.ptr, 0xFF, dQ.nbytes, st.handle), "vj_memset")
        d = DeviceArray.from_host(padded(img, stride))
        check(lib().vj_integral_image_noalloc(d.ptr, w, h, stride, dS.ptr, dQ.ptr, st.handle),
              "vj_integral_image_noalloc")
        st.synchronize()
        gs = dS.to_host()[:n].reshape(h + 1, w + 1)
        gq = dQ.to_host()[:n].reshape(h + 1, w + 1)
        es, eq = via_engine(img, stride)
        assert np.array_equal(gs, es) and np.array_equal(gq, eq)
        os_, oq = O.integral(img)
        assert np.array_equal(gs, os_) and np.array_equal(gq, oq)
        # the rest of the caller's buffer is untouched
        assert (dS.to_host()[n:] == 0xFFFFFFFF).all()


def test_noalloc_integral_refuses_oversized_images():
    st = Stream()
    d = DeviceArray(64, np.uint8)
    out = DeviceArray(64, np.uint64)
    for w, h in [(8193, 10), (10, 8193), (8192, 1029)]:  # the last: 8192 * 1029 * 255 >= 2^31
        assert lib().vj_integral_image_noalloc(d.ptr, w, h, w, out.ptr, out.ptr, st.handle) == VJ_INVALID_ARG
    assert lib().vj_integral_image_noalloc(d.ptr, 8, 8, 7, out.ptr, out.ptr, st.handle) == VJ_INVALID_ARG
