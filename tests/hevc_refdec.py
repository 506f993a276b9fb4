"""The reference's own H.265 decoding functions as a test oracle (test
helper, not a test module): oracle/_ref/libref_hevc_dec.so, built by
`make -C oracle ref` from the reference's src/hevc_ps.c, hevc_cu.c,
hevc_mv.c, hevc_residual.c, hevc_pred.c, hevc_transform.c, hevc_mc.c,
hevc_filter.c (with their tables and thread pool) and the harness
oracle/ref_hevc_decode.c, which walks an Annex B stream in place of the
reference's src/decoder_h265.c (not buildable here: it includes Vulkan).
Absent when the reference tree was not there to build it."""
import ctypes as C
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "_ref", "libref_hevc_dec.so")
_lib = None


def available():
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB)
        _lib.ref_hevc_decode.restype = C.c_int
        _lib.ref_hevc_decode.argtypes = [C.c_char_p, C.c_size_t, C.c_void_p, C.c_size_t, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]
        for f in (_lib.ref_hevc_read_sps, _lib.ref_hevc_read_pps):
            f.restype = C.c_int
            f.argtypes = [C.c_char_p, C.c_size_t, C.c_void_p]
    return _lib


def decode(stream, width, height, bit_depth, npics):
    """Pictures of an Annex B stream as [(y, cb, cr)] at the coded size
    (uint8, or uint16 samples above 8 bits)."""
    bps = 2 if bit_depth > 8 else 1
    per = width * height * bps * 3 // 2
    out = np.zeros(per * npics, np.uint8)
    w, h, bd = C.c_int(0), C.c_int(0), C.c_int(0)
    n = lib().ref_hevc_decode(stream, len(stream), out.ctypes.data, out.size, C.byref(w), C.byref(h), C.byref(bd))
    if n < 0:
        raise RuntimeError(f"reference decoder refused the stream ({n})")
    assert (w.value, h.value, bd.value) == (width, height, bit_depth)
    pics = []
    dt = np.uint16 if bps == 2 else np.uint8
    for i in range(n):
        a = out[i * per:(i + 1) * per].view(dt)
        ny, nc = width * height, width * height // 4
        pics.append((a[:ny].reshape(height, width), a[ny:ny + nc].reshape(height // 2, width // 2),
                     a[ny + nc:].reshape(height // 2, width // 2)))
    return pics


def nal_units(stream):
    """(type, payload with header, without start code) of each NAL unit."""
    out, i, n = [], 0, len(stream)
    starts = []
    while i + 3 <= n:
        if stream[i] == 0 and stream[i + 1] == 0 and stream[i + 2] == 1:
            starts.append(i + 3)
            i += 3
        else:
            i += 1
    for k, a in enumerate(starts):
        b = starts[k + 1] - 3 if k + 1 < len(starts) else n
        while b > a and stream[b - 1] == 0:
            b -= 1
        out.append(((stream[a] >> 1) & 63, stream[a:b]))
    return out


SPS_FIELDS = ("width", "height", "crop_left", "crop_right", "crop_top", "crop_bottom", "bit_depth_luma",
              "bit_depth_chroma", "log2_max_poc_lsb", "log2_min_cb", "log2_ctb", "log2_min_tb", "log2_max_tb",
              "max_th_depth_inter", "max_th_depth_intra", "num_st_rps", "temporal_mvp", "amp", "sao",
              "chroma_format_idc")
PPS_FIELDS = ("init_qp", "deblocking_disabled", "loop_filter_across_slices", "cu_qp_delta", "sign_data_hiding",
              "entropy_coding_sync", "tiles", "num_ref_idx_l0_default", "transform_skip", "dependent_slices")


def read_sps(nal):
    f = (C.c_int32 * len(SPS_FIELDS))()
    rc = lib().ref_hevc_read_sps(nal, len(nal), f)
    if rc:
        raise RuntimeError(f"reference SPS parser refused ({rc})")
    return dict(zip(SPS_FIELDS, list(f)))


def read_pps(nal):
    f = (C.c_int32 * len(PPS_FIELDS))()
    rc = lib().ref_hevc_read_pps(nal, len(nal), f)
    if rc:
        raise RuntimeError(f"reference PPS parser refused ({rc})")
    return dict(zip(PPS_FIELDS, list(f)))
