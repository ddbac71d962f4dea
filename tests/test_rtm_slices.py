"""RTM slice capture and the slice frame format: the GPU gather
(stp_rtm_extract_slice_f64 through RTMEngine.extract_slices) against the
reference's RTMEngine::extract_slices compiled from its own source
(oracle/_ref), bit-exact; the frame codec against hand-computed checksums and
the reference's field layout."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from seismic_engine.rtm import (SLICE_HEADER_SIZE, CapturedSlice, SliceAxis, SliceSpec, decode_slice_frame,
                                encode_slice_frame, slice_checksum)

ROOT = Path(__file__).resolve().parent.parent
REFLIB = ROOT / "oracle" / "_ref" / "libref.so"
N = (9, 7, 11)
HW = 3


def test_checksum_hand_values():
    words = np.array([1, 2], "<u4").tobytes()
    assert slice_checksum(words) == 0x4004000          # ((1 rotl 13) ^ 2) rotl 13
    assert slice_checksum(words + b"\x05") == 0x4004005  # trailing byte xor-ed without rotation
    assert slice_checksum(b"") == 0


def test_frame_roundtrip_layout_and_corruption():
    rng = np.random.default_rng(0)
    sl = CapturedSlice(17, SliceAxis.Y, 3, 4, 5, rng.standard_normal(20), rng.standard_normal(20),
                       rng.standard_normal(20), 123.5)
    buf = encode_slice_frame(sl, sequence_id=9)
    assert len(buf) == SLICE_HEADER_SIZE + 3 * 20 * 8
    # the reference's packed field offsets (snapshot_buffer.h:18-37)
    assert int.from_bytes(buf[0:8], "little") == 0x534C494345303031
    assert int.from_bytes(buf[16:20], "little") == 1
    assert int.from_bytes(buf[44:52], "little") == len(buf)
    assert [int.from_bytes(buf[o:o + 8], "little") for o in (52, 60, 68)] == [128, 288, 448]
    assert buf[124:128] == b"\x00" * 4
    out, seq = decode_slice_frame(buf)
    assert seq == 9 and out.timestep == 17 and out.axis == SliceAxis.Y and (out.dim0, out.dim1) == (4, 5)
    for a, b in ((out.forward_slice, sl.forward_slice), (out.adjoint_slice, sl.adjoint_slice),
                 (out.ic_slice, sl.ic_slice)):
        assert np.array_equal(a, b)
    bad = bytearray(buf)
    bad[200] ^= 1
    assert decode_slice_frame(bytes(bad)) is None
    bad = bytearray(buf)
    bad[0] ^= 1
    assert decode_slice_frame(bytes(bad)) is None
    assert decode_slice_frame(buf[:100]) is None


def _ref_slices(fwd, adj, ic, specs, HW=HW):
    lib = C.CDLL(str(REFLIB))
    lib.ref_rtm_create.restype = C.c_void_p
    h = C.c_void_p(lib.ref_rtm_create(C.c_int64(N[0]), C.c_int64(N[1]), C.c_int64(N[2]), C.c_int(HW)))
    axes = np.array([int(s.axis) for s in specs], np.int32)
    idx = np.array([s.index for s in specs], np.int64)
    out = np.zeros(3 * 3 * max(N) ** 2 * len(specs))
    dims = np.zeros(2 * len(specs), np.int64)
    nyh, nzh = N[1] + 2 * HW, N[2] + 2 * HW
    cnt = lib.ref_rtm_extract(h, C.c_int(len(specs)), axes.ctypes.data_as(C.c_void_p),
                              idx.ctypes.data_as(C.c_void_p), fwd.ctypes.data_as(C.c_void_p),
                              adj.ctypes.data_as(C.c_void_p), ic.ctypes.data_as(C.c_void_p), C.c_int64(nyh),
                              C.c_int64(nzh), C.c_int64(5), out.ctypes.data_as(C.c_void_p),
                              dims.ctypes.data_as(C.c_void_p))
    lib.ref_rtm_free(h)
    res, off = [], 0
    for k in range(cnt):
        n = dims[2 * k] * dims[2 * k + 1]
        res.append((dims[2 * k], dims[2 * k + 1], out[off:off + n], out[off + n:off + 2 * n],
                    out[off + 2 * n:off + 3 * n]))
        off += 3 * n
    return res


@pytest.mark.gpu
@pytest.mark.skipif(not REFLIB.exists(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("HW", [0, 3, 4])
def test_gpu_slices_match_reference(HW, cuda):
    from seismic_engine.rtm import RTMEngine
    rng = np.random.default_rng(1)
    nh = (N[0] + 2 * HW) * (N[1] + 2 * HW) * (N[2] + 2 * HW)
    fwd = rng.standard_normal(nh)
    adj = rng.standard_normal(nh)
    ic = rng.standard_normal(N[0] * N[1] * N[2])
    fwd[[5, nh // 2, nh - 7]] = [np.nan, np.inf, -np.inf]
    adj[[7, 401]] = np.nan
    ic[[0, 3]] = np.nan
    specs = [SliceSpec(SliceAxis.X, 2), SliceSpec(SliceAxis.Y, -1), SliceSpec(SliceAxis.Z, 10),
             SliceSpec(SliceAxis.VOLUME, -1), SliceSpec(SliceAxis.Z, -1)]
    ref = _ref_slices(fwd, adj, ic, specs, HW)
    e = RTMEngine(*N, halo_width=HW)
    e.set_slice_specs(specs)
    got = []
    e.set_slice_callback(got.append)
    e.extract_slices(fwd, adj, ic, N[1] + 2 * HW, N[2] + 2 * HW, 5)
    assert e.total_captured_slices() == len(ref) == 4
    for s, (d0, d1, f, a, c) in zip(got, ref):
        assert (s.dim0, s.dim1) == (d0, d1) and s.timestep == 5
        assert np.array_equal(s.forward_slice, f) and np.array_equal(s.adjoint_slice, a)
        assert np.array_equal(s.ic_slice, c)
        assert decode_slice_frame(encode_slice_frame(s))[0].index == s.index


@pytest.mark.gpu
@pytest.mark.skipif(not REFLIB.exists(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("HW", [0, 4])
def test_gpu_image_condition_and_finalize_match_reference(HW, cuda):
    """stp_rtm_image_condition_f64 / stp_rtm_finalize_f64 against the
    reference's RTMEngine::apply_image_condition, finalize_image and
    reset_image (rtm_engine.cpp:24-57, 170-190) compiled from its own source:
    three accumulations (non-finite inputs included), finalize, then a reset
    and one more accumulation."""
    import torch
    from seismic_engine._lib import call, ptr, stream_handle
    lib = C.CDLL(str(REFLIB))
    lib.ref_rtm_create.restype = C.c_void_p
    for f in ("ref_rtm_apply", "ref_rtm_finalize", "ref_rtm_reset", "ref_rtm_result", "ref_rtm_free"):
        getattr(lib, f).restype = None
    lib.ref_rtm_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64]
    lib.ref_rtm_finalize.argtypes = lib.ref_rtm_reset.argtypes = lib.ref_rtm_free.argtypes = [C.c_void_p]
    lib.ref_rtm_result.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    nyh, nzh = N[1] + 2 * HW, N[2] + 2 * HW
    nh = (N[0] + 2 * HW) * nyh * nzh
    ni = N[0] * N[1] * N[2]
    rng = np.random.default_rng(7)
    h = C.c_void_p(lib.ref_rtm_create(C.c_int64(N[0]), C.c_int64(N[1]), C.c_int64(N[2]), C.c_int(HW)))
    dev = dict(dtype=torch.float64, device=cuda)
    img, ill, refl = (torch.zeros(ni, **dev) for _ in range(3))

    def ref_result():
        out = [np.zeros(ni) for _ in range(3)]
        lib.ref_rtm_result(h, *(o.ctypes.data_as(C.c_void_p) for o in out))
        return out

    def both(f, a, step):
        lib.ref_rtm_apply(h, f.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p), nyh, nzh, step)
        ft, at = torch.as_tensor(f, **dev), torch.as_tensor(a, **dev)
        call("stp_rtm_image_condition_f64", C.c_int64(N[0]), C.c_int64(N[1]), C.c_int64(N[2]), HW,
             C.c_int64(nyh), C.c_int64(nzh), ptr(ft), ptr(at), ptr(img), ptr(ill), stream_handle())

    for step in range(3):
        f, a = rng.standard_normal(nh), rng.standard_normal(nh)
        f[[3 + step, nh // 3]] = [np.nan, np.inf]
        a[[11, nh - 2]] = [-np.inf, np.nan]
        both(f, a, step)
    # a cell with zero illumination: finalize must give 0 there
    lib.ref_rtm_finalize(h)
    call("stp_rtm_finalize_f64", C.c_int64(ni), ptr(img), ptr(ill), ptr(refl), stream_handle())
    r_img, r_ill, r_refl = ref_result()
    np.testing.assert_allclose(img.cpu().numpy(), r_img, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(ill.cpu().numpy(), r_ill, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(refl.cpu().numpy(), r_refl, rtol=1e-13, atol=1e-13)
    # reset, then one accumulation of a field with itself (f = a)
    lib.ref_rtm_reset(h)
    img.zero_(); ill.zero_()
    f = rng.standard_normal(nh)
    both(f, f, 9)
    lib.ref_rtm_finalize(h)
    call("stp_rtm_finalize_f64", C.c_int64(ni), ptr(img), ptr(ill), ptr(refl), stream_handle())
    r_img, r_ill, r_refl = ref_result()
    lib.ref_rtm_free(h)
    np.testing.assert_allclose(img.cpu().numpy(), r_img, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(refl.cpu().numpy(), r_refl, rtol=1e-13, atol=1e-13)
