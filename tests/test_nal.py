"""NAL packaging of the slices (host/h264_host.c): emulation prevention
against the reference's own bs_rbsp_to_ebsp (src/bitstream.c:232-250,
compiled into oracle/_ref/libref_cavlc.so), its defining properties, and the
slice NAL layout of src/encoder_h264.c:3440-3470."""
import ctypes as C
import os

import numpy as np
import pytest

import orc
from conftest import load_pkg


@pytest.fixture(scope="module")
def pkg():
    return load_pkg()


def unescape(ebsp):
    out, zeros, i = bytearray(), 0, 0
    while i < len(ebsp):
        b = ebsp[i]
        if zeros >= 2 and b == 3:
            zeros, i = 0, i + 1
            continue
        out.append(b)
        zeros = zeros + 1 if b == 0 else 0
        i += 1
    return bytes(out)


def zero_rich(rs, n):
    a = rs.integers(0, 256, n).astype(np.uint8)
    a[rs.random(n) < 0.6] = 0
    a[rs.random(n) < 0.1] = rs.integers(1, 4)
    return a.tobytes()


def test_ebsp_properties(pkg):
    rs = np.random.default_rng(3)
    for n in list(range(0, 12)) + [100, 1000, 5000]:
        for _ in range(20):
            r = zero_rich(rs, n)
            e = pkg.rbsp_to_ebsp(r)
            assert unescape(e) == r
            # 7.4.1: 00 00 00, 00 00 01 and 00 00 02 never occur; 00 00 03 is the escape
            for i in range(len(e) - 2):
                assert not (e[i] == 0 and e[i + 1] == 0 and e[i + 2] <= 2)
    assert pkg.rbsp_to_ebsp(b"\x00\x00\x00\x00\x01") == b"\x00\x00\x03\x00\x00\x03\x01"
    assert pkg.rbsp_to_ebsp(b"\x00\x00\x04") == b"\x00\x00\x04"


@pytest.mark.skipif(not os.path.exists(orc.REF_CAVLC), reason="oracle/_ref not built (reference tree absent)")
def test_ebsp_matches_reference(pkg):
    ref = C.CDLL(orc.REF_CAVLC)
    ref.bs_rbsp_to_ebsp.restype = C.c_size_t
    ref.bs_rbsp_to_ebsp.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
    rs = np.random.default_rng(4)
    for n in [1, 2, 3, 7, 64, 777, 4096]:
        for _ in range(30):
            r = np.frombuffer(zero_rich(rs, n), np.uint8)
            d = np.zeros(2 * n + 4, np.uint8)
            k = ref.bs_rbsp_to_ebsp(d.ctypes.data_as(C.c_void_p), d.size, r.ctypes.data_as(C.c_void_p), n)
            assert pkg.rbsp_to_ebsp(r.tobytes()) == d[:k].tobytes()


def test_frame_nals_layout(pkg):
    rs = np.random.default_rng(5)
    from test_oracle_cavlc import synth_frame
    for idr in (1, 0):
        lv, dc, modes, mvs = synth_frame(rs, 6, 4, idr)
        slices = orc.h264_cavlc_slices(lv, dc, modes, mvs, orc.slice_cfg(6, 4, 27, idr, 3))
        out = pkg.h264_frame_nals(slices, idr)
        parts = out.split(b"\x00\x00\x00\x01")
        assert parts[0] == b"" and len(parts) == 4
        for p, s in zip(parts[1:], slices):
            assert p[0] == (0x65 if idr else 0x41)
            assert unescape(p[1:]) == s


class _RefSps(C.Structure):  # the reference's h264_sps_t (bitstream.h), natural alignment
    _fields_ = [(n, C.c_uint8) for n in ("profile_idc", "level_idc", "sps_id", "chroma_format_idc", "bit_depth_luma",
                                         "bit_depth_chroma", "log2_max_frame_num", "pic_order_cnt_type",
                                         "log2_max_poc_lsb")] + \
               [(n, C.c_uint32) for n in ("max_num_ref_frames", "pic_width_in_mbs", "pic_height_in_mbs")] + \
               [(n, C.c_bool) for n in ("frame_mbs_only", "direct_8x8_inference", "frame_cropping")] + \
               [(n, C.c_uint32) for n in ("crop_left", "crop_right", "crop_top", "crop_bottom")] + \
               [("vui_present", C.c_bool), ("sar_width", C.c_uint32), ("sar_height", C.c_uint32),
                ("timing_info_present", C.c_bool), ("num_units_in_tick", C.c_uint32), ("time_scale", C.c_uint32)]


class _RefPps(C.Structure):  # the reference's h264_pps_t
    _fields_ = [("pps_id", C.c_uint8), ("sps_id", C.c_uint8), ("entropy_coding_mode", C.c_bool),
                ("pic_order_present", C.c_bool), ("num_ref_idx_l0_default", C.c_uint8),
                ("num_ref_idx_l1_default", C.c_uint8), ("weighted_pred", C.c_bool), ("weighted_bipred_idc", C.c_uint8),
                ("pic_init_qp", C.c_int8), ("chroma_qp_offset", C.c_int8), ("second_chroma_qp_offset", C.c_int8),
                ("deblocking_filter_control", C.c_bool), ("constrained_intra_pred", C.c_bool),
                ("transform_8x8_mode", C.c_bool)]


SPS_FIELDS = ("profile_idc", "level_idc", "sps_id", "chroma_format_idc", "bit_depth_luma", "bit_depth_chroma",
              "log2_max_frame_num", "pic_order_cnt_type", "log2_max_poc_lsb", "max_num_ref_frames", "pic_width_in_mbs",
              "pic_height_in_mbs", "frame_mbs_only", "direct_8x8_inference", "frame_cropping", "crop_left",
              "crop_right", "crop_top", "crop_bottom", "vui_present", "sar_width", "sar_height",
              "timing_info_present", "num_units_in_tick", "time_scale")
PPS_FIELDS = ("pps_id", "sps_id", "entropy_coding_mode", "pic_order_present", "num_ref_idx_l0_default",
              "num_ref_idx_l1_default", "weighted_pred", "weighted_bipred_idc", "pic_init_qp", "chroma_qp_offset",
              "second_chroma_qp_offset", "deblocking_filter_control", "constrained_intra_pred", "transform_8x8_mode")


@pytest.mark.skipif(not os.path.exists(orc.REF_CAVLC), reason="oracle/_ref not built (reference tree absent)")
def test_parameter_sets_match_reference(pkg):
    ref = C.CDLL(orc.REF_CAVLC)
    ref.h264_sps_default.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint8]
    ref.h264_pps_default.argtypes = [C.c_void_p, C.c_uint8, C.c_bool, C.c_int32]
    for f in ("bs_write_sps", "bs_write_pps"):
        getattr(ref, f).restype = C.c_size_t
        getattr(ref, f).argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
    for f in ("bc250_h264_write_sps", "bc250_h264_write_pps"):
        getattr(pkg.host, f).restype = C.c_size_t
        getattr(pkg.host, f).argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    a, b = np.zeros(512, np.uint8), np.zeros(512, np.uint8)
    pa, pb = a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)
    cases = 0
    for (w, h) in ((1920, 1080), (1280, 720), (640, 480), (200, 120), (3840, 2160)):
        for fps in (30, 60):
            for prof in (66, 77, 100):
                for vui in (1, 0):
                    s = _RefSps()
                    ref.h264_sps_default(C.byref(s), w, h, fps, prof)
                    s.vui_present = bool(vui)
                    s.timing_info_present = fps == 60 or not vui
                    mine = np.array([int(getattr(s, k)) for k in SPS_FIELDS], np.int32)
                    n1 = ref.bs_write_sps(pa, 512, C.byref(s))
                    n2 = pkg.host.bc250_h264_write_sps(mine.ctypes.data_as(C.c_void_p), pb, 512)
                    assert n1 == n2 and (a[:n1] == b[:n2]).all(), (w, h, fps, prof, vui)
                    cases += 1
    for cabac in (False, True):
        for qp in (0, 22, 26, 51):
            for t8 in (0, 1):
                s = _RefPps()
                ref.h264_pps_default(C.byref(s), 0, cabac, qp)
                s.transform_8x8_mode = bool(t8)
                s.second_chroma_qp_offset = -2 if t8 else 0
                mine = np.array([int(getattr(s, k)) for k in PPS_FIELDS], np.int32)
                n1 = ref.bs_write_pps(pa, 512, C.byref(s))
                n2 = pkg.host.bc250_h264_write_pps(mine.ctypes.data_as(C.c_void_p), pb, 512)
                assert n1 == n2 and (a[:n1] == b[:n2]).all(), (cabac, qp, t8)
                cases += 1
    assert cases == 76


def _ps_len(pkg, fields, vals, which):
    f = pkg.host.bc250_h264_write_sps if which == "sps" else pkg.host.bc250_h264_write_pps
    f.restype = C.c_size_t
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    a = np.array([int(vals.get(k, 0)) for k in fields], np.int32)
    buf = np.zeros(512, np.uint8)
    return f(a.ctypes.data_as(C.c_void_p), buf.ctypes.data_as(C.c_void_p), 512)


def test_parameter_sets_refuse_out_of_range_fields(pkg):
    """Fields outside their 7.4.2.1.1 / 7.4.2.2 ranges give no parameter set
    (0 bytes) instead of an oversized Exp-Golomb code; the range ends are
    accepted."""
    base_sps = dict(profile_idc=100, level_idc=40, sps_id=0, chroma_format_idc=1, bit_depth_luma=8,
                    bit_depth_chroma=8, log2_max_frame_num_minus4=0, pic_order_cnt_type=0,
                    log2_max_poc_lsb_minus4=2, max_num_ref_frames=1, width_in_mbs=120, height_in_mbs=68,
                    frame_mbs_only=1, direct_8x8_inference=1, frame_cropping=1, crop_bottom=4, vui_present=1,
                    sar_width=1, sar_height=1, timing_info_present=1, num_units_in_tick=1, time_scale=60)
    assert _ps_len(pkg, pkg.SPS_FIELDS, base_sps, "sps") > 0
    good = [("sps_id", 31), ("bit_depth_luma", 14), ("log2_max_frame_num_minus4", 12), ("pic_order_cnt_type", 2),
            ("log2_max_poc_lsb_minus4", 12), ("max_num_ref_frames", 16), ("width_in_mbs", 1), ("crop_bottom", 0),
            ("sar_width", 65535), ("chroma_format_idc", 3)]
    bad = [("sps_id", 32), ("sps_id", -1), ("bit_depth_luma", 7), ("bit_depth_chroma", 15),
           ("chroma_format_idc", 4), ("log2_max_frame_num_minus4", 13), ("log2_max_frame_num_minus4", -1),
           ("pic_order_cnt_type", 1), ("pic_order_cnt_type", 3), ("log2_max_poc_lsb_minus4", 13),
           ("max_num_ref_frames", 17), ("width_in_mbs", 0), ("height_in_mbs", 0), ("crop_bottom", -1),
           ("crop_bottom", 544), ("crop_right", 960), ("sar_width", 65536), ("num_units_in_tick", 0),
           ("time_scale", -5), ("profile_idc", 256), ("level_idc", -1)]
    for k, v in good:
        assert _ps_len(pkg, pkg.SPS_FIELDS, dict(base_sps, **{k: v}), "sps") > 0, (k, v)
    for k, v in bad:
        assert _ps_len(pkg, pkg.SPS_FIELDS, dict(base_sps, **{k: v}), "sps") == 0, (k, v)
    # fields a profile does not write are not checked
    assert _ps_len(pkg, pkg.SPS_FIELDS, dict(base_sps, profile_idc=66, bit_depth_luma=0, chroma_format_idc=9), "sps") > 0
    base_pps = dict(pps_id=0, sps_id=0, num_ref_idx_l0_default=0, pic_init_qp_minus26=0, deblocking_filter_control=1,
                    transform_8x8_mode=1, second_chroma_qp_offset=0)
    assert _ps_len(pkg, pkg.PPS_FIELDS, base_pps, "pps") > 0
    for k, v in [("pps_id", 255), ("num_ref_idx_l0_default", 31), ("pic_init_qp_minus26", -26),
                 ("pic_init_qp_minus26", 25), ("chroma_qp_offset", -12), ("second_chroma_qp_offset", 12),
                 ("weighted_bipred_idc", 2)]:
        assert _ps_len(pkg, pkg.PPS_FIELDS, dict(base_pps, **{k: v}), "pps") > 0, (k, v)
    for k, v in [("pps_id", 256), ("sps_id", 32), ("num_ref_idx_l0_default", 32), ("num_ref_idx_l1_default", -1),
                 ("weighted_bipred_idc", 3), ("pic_init_qp_minus26", 26), ("pic_init_qp_minus26", -63),
                 ("chroma_qp_offset", 13), ("second_chroma_qp_offset", -13)]:
        assert _ps_len(pkg, pkg.PPS_FIELDS, dict(base_pps, **{k: v}), "pps") == 0, (k, v)
