"""Python access to the MI355X encoder pipeline (libbc250_hip.so) and its
host C library (libbc250_host.so), through the C ABI of include/gpu_compute.h
and include/bc250_stages.h.

The libraries are built in-tree (``make -C bc250-vaapi_amd``); importing this
module without them raises, there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(_HERE, "libbc250_hip.so")
HOST_LIB = os.path.join(_HERE, "libbc250_host.so")

GPU_IMAGE_NV12 = 0
GPU_CODEC_H264, GPU_CODEC_HEVC = 0, 1
GPU_IMAGE_P010 = 1
# forms of the 32x32 transform batches (include/bc250_hevc.h)
T32_BUTTERFLY, T32_MFMA, T32_MATRIX = 0, 1, 2


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run `make -C {_HERE}` (the HIP extension is required; there is no CPU path)")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


hip = _load(HIP_LIB)
host = _load(HOST_LIB)


class GpuImage(C.Structure):
    """gpu_image_t (include/gpu_compute.h)."""
    _fields_ = [("y_plane", C.c_void_p), ("uv_plane", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32), ("coded_width", C.c_uint32), ("coded_height", C.c_uint32),
                ("y_pitch", C.c_uint32), ("uv_pitch", C.c_uint32), ("current_layout", C.c_int),
                ("format", C.c_int)]


class GpuMemory(C.Structure):
    _fields_ = [("memory", C.c_void_p), ("size", C.c_size_t), ("mapped_ptr", C.c_void_p)]


class Nv12Layout(C.Structure):
    """gpu_nv12_layout_t (include/gpu_compute.h)."""
    _fields_ = [("y_pitch", C.c_uint32), ("y_offset", C.c_uint64), ("uv_pitch", C.c_uint32),
                ("uv_offset", C.c_uint64), ("total_size", C.c_uint64)]


class GpuContext(C.Structure):
    _fields_ = [("hip", C.c_void_p), ("device", C.c_int), ("frames_submitted", C.c_uint32),
                ("frame_width", C.c_uint32), ("frame_height", C.c_uint32),
                ("recon_image", GpuImage), ("recon_memory", GpuMemory),
                ("has_recon_frame", C.c_bool), ("any_surface_exported", C.c_bool)]


class EncodeParams(C.Structure):
    _fields_ = [("search_range", C.c_int), ("lambda8", C.c_int), ("deblock_idc", C.c_int),
                ("me_8x8", C.c_int), ("readback", C.c_int), ("codec", C.c_int), ("square_radius", C.c_int),
                ("diamond", C.c_int), ("hevc_ctu_log2", C.c_int), ("entropy", C.c_int)]


class CpuSimdMeConfig(C.Structure):
    """cpu_simd_me_config_t (bc250_cpu_me.h)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("width_in_mbs", C.c_uint32),
                ("height_in_mbs", C.c_uint32), ("search_radius", C.c_uint32), ("num_threads", C.c_int),
                ("core_ids", C.c_int * 2)]


class StageStats(C.Structure):
    _fields_ = [("sum_ms", C.c_double * 5), ("total_ms", C.c_double), ("frames", C.c_int), ("p_frames", C.c_int),
                ("me_int_ms_p", C.c_double)]


class SeqCfg(C.Structure):
    _fields_ = [("qp", C.c_int), ("gop", C.c_int), ("num_slices", C.c_int), ("pipelined", C.c_int)]


MV_DTYPE = np.dtype([("mvx", "<i4"), ("mvy", "<i4"), ("sad", "<u4"), ("pad", "<u4")])

_P = C.c_void_p
_I = C.c_int
_CTX = C.POINTER(GpuContext)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig(hip, "bc250_gpu_init", _I, [_CTX])
_sig(hip, "bc250_gpu_destroy", None, [_CTX])
_sig(hip, "gpu_compute_create_image", _I, [_CTX, _I, _I, _I, C.POINTER(GpuImage), C.POINTER(GpuMemory)])
_sig(hip, "gpu_compute_destroy_image", None, [_CTX, GpuImage, GpuMemory])
_sig(hip, "gpu_compute_upload_nv12", _I, [_CTX, C.POINTER(GpuImage), GpuMemory, _P, _I, _P, _I, _I, _I])
_sig(hip, "gpu_compute_upload_p010", _I, [_CTX, C.POINTER(GpuImage), GpuMemory, _P, _I, _P, _I, _I, _I])
_sig(hip, "gpu_compute_download_nv12", _I, [_CTX, C.POINTER(GpuImage), GpuMemory, _P, _I, _P, _I, _I, _I])
_sig(hip, "gpu_compute_begin_picture", _I, [_CTX, GpuImage])
_sig(hip, "gpu_compute_dispatch_encode_ext", _I, [_CTX, GpuImage, _I, _I, _I, _I, _I, _I, _P])
_sig(hip, "gpu_compute_dispatch_me_only", _I, [_CTX, GpuImage, _I, _I])
_sig(hip, "gpu_compute_end_picture", _I, [_CTX])
_sig(hip, "gpu_compute_submitted_slot", _I, [_CTX])
_sig(hip, "gpu_compute_sync_slot", _I, [_CTX, _I])
for _n in ("quant", "dc", "pred_mode", "mv", "nz", "mv8", "hevc_mode"):
    _sig(hip, f"gpu_compute_get_{_n}_staging_data_slot", _I, [_CTX, _I, C.POINTER(_P), C.POINTER(C.c_size_t)])
_sig(hip, "gpu_compute_set_encode_params", _I, [_CTX, C.POINTER(EncodeParams)])
_sig(hip, "gpu_compute_download_recon", _I, [_CTX, _P, _I, _P, _I])
_sig(hip, "gpu_compute_get_stage_times", _I, [_CTX, _I, C.POINTER(C.c_float)])
_sig(hip, "gpu_compute_set_stage_timing", _I, [_CTX, _I])
_sig(hip, "gpu_compute_take_error_flags", _I, [_CTX])
_sig(hip, "gpu_compute_get_stage_stats", _I, [_CTX, C.POINTER(StageStats), _I])
_sig(hip, "gpu_compute_get_last_latency_ms", C.c_double, [_CTX])
_sig(hip, "gpu_compute_cpu_read_safe", _I, [_CTX])
_sig(hip, "gpu_compute_wait_for_image_ready_host", _I, [_CTX, GpuMemory])
_sig(hip, "gpu_compute_get_nv12_layout", _I, [_CTX, C.POINTER(GpuImage), GpuMemory, C.POINTER(Nv12Layout)])
_sig(hip, "gpu_compute_map_surface", _P, [_CTX, C.POINTER(GpuImage), GpuMemory, C.POINTER(Nv12Layout),
                                          C.POINTER(C.c_bool)])
_sig(hip, "gpu_compute_unmap_surface", None, [_CTX, GpuMemory, C.c_bool])

_U = C.c_uint
_sig(hip, "bc250_launch_me_fullsearch", _I, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _P, _P])
_sig(hip, "bc250_launch_me_fullsearch_list", _I, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P])
_sig(hip, "bc250_launch_me_fullsearch_list_split", _I, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _I,
                                                        _P])
# the learned proposer (include/bc250_learned.h)
_F = C.c_float
_sig(hip, "bc250_cnn_param_count", C.c_size_t, [])
_sig(hip, "bc250_cnn_input", _I, [_P, _I, _P, _I, C.c_size_t, _I, _I, _P, _I, _P, _P])
_sig(hip, "bc250_cnn_conv", _I, [_I, _I, _P, _P, _P, _P, _P, _P, _I, _P])
_sig(hip, "bc250_cnn_pool", _I, [_P, _P, _I, _P])
_sig(hip, "bc250_cnn_forward", _I, [_P, _P, _P, _P, _I, _P])
_sig(hip, "bc250_cnn_forward_picture", _I, [_P, _I, _P, _I, _I, _I, _P, _P, _P, _P])
_sig(hip, "bc250_cnn_forward_propose_picture", _I, [_P, _I, _P, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P])
_sig(hip, "bc250_cnn_set_forward_variant", _I, [_I])
_sig(hip, "bc250_cnn_wgrad", _I, [_I, _P, _P, _P, _P, _I, _P])
_sig(hip, "bc250_cnn_head", _I, [_P, _P, _P, _P, _I, _P])
_sig(hip, "bc250_cnn_head_train", _I, [_P, _P, _P, _P, _P, _P, _P, _P, _I, _F, _P])
_sig(hip, "bc250_cnn_pool_backward", _I, [_P, _P, _P, _I, _P])
_sig(hip, "bc250_cnn_adam", _I, [_P, _P, _P, _P, _I, _F, _I, _P])
_sig(hip, "bc250_cnn_pack", _I, [_P, _P, _P])
_sig(hip, "bc250_cnn_mvin", _I, [_P, _I, _I, _P, _P])
_sig(hip, "bc250_cnn_candidates", _I, [_P, _I, _I, _I, _P, _P])
_sig(hip, "bc250_launch_cand_search", _I, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _I, _P, _P, _P])
_sig(hip, "bc250_fallback_select", _I, [_P, _I, _I, _I, _I, _I, _P, _P, _P])
_sig(hip, "bc250_cnn_propose", _I, [_P, _P, _P, _I, _I, _I, _P, _P])
_sig(hip, "bc250_learned_search_picture", _I, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _I, _P, _P,
                                               _P, _P, _P, _P, _P, _P])


class ProposerCfg(C.Structure):
    """bc250_proposer_cfg_t (include/bc250_learned.h)."""
    _fields_ = [("ncand", C.c_int32), ("alpha_q4", C.c_int32), ("beta", C.c_int32), ("t_high", C.c_int32)]


_sig(hip, "gpu_compute_set_me_proposer", _I, [_CTX, _P, C.POINTER(ProposerCfg)])
_sig(hip, "bc250_cnn_set_grid_cap", _I, [_I])


# ---------------------------------------------------------------- serving (include/bc250_serve.h)

class ServerCfg(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_batch", C.c_int32), ("learned", C.c_int32),
                ("search_range", C.c_int32), ("lambda8", C.c_int32), ("proposer", ProposerCfg)]


class SessionCfg(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("codec", C.c_int32), ("qp", C.c_int32),
                ("num_slices", C.c_int32), ("gop", C.c_int32), ("ring", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("session", C.c_int32), ("frame", C.c_int32), ("source", C.c_int32), ("pad", C.c_int32),
                ("t_arrival", C.c_double)]


class FrameResult(C.Structure):
    _fields_ = [("session", C.c_int32), ("frame", C.c_int32), ("batch", C.c_int32), ("fallbacks", C.c_int32),
                ("t_arrival", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("checksum", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("batches", C.c_int32), ("frames", C.c_int32), ("max_batch_seen", C.c_int32),
                ("sessions", C.c_int32), ("mean_batch", C.c_double), ("forward_variant", C.c_int32),
                ("autotune_us", C.c_double * 8), ("hbm_used", C.c_size_t)]


_sig(hip, "bc250_server_create", _P, [C.POINTER(ServerCfg), _P])
_sig(hip, "bc250_server_destroy", None, [_P])
_sig(hip, "bc250_server_session_bytes", C.c_size_t, [_P, C.POINTER(SessionCfg)])
_sig(hip, "bc250_server_open", _I, [_P, C.POINTER(SessionCfg)])
_sig(hip, "bc250_server_close", _I, [_P, _I])
_sig(hip, "bc250_server_upload", _I, [_P, _I, _I, _P, _I, _P, _I])
_sig(hip, "bc250_server_run", _I, [_P, C.POINTER(Request), _I, _I, C.POINTER(FrameResult), C.POINTER(C.c_double)])
_sig(hip, "bc250_server_autotune", _I, [_P])
_sig(hip, "bc250_server_stats", _I, [_P, C.POINTER(ServerStats)])
_sig(hip, "bc250_fnv1a64", C.c_uint64, [C.c_uint64, _P, C.c_size_t])


class Server:
    """Many sessions on the GPU (include/bc250_serve.h)."""

    def __init__(self, learned=False, max_batch=32, hbm_budget=0, search_range=64, lambda8=1, params=None,
                 proposer=None):
        pc = ProposerCfg(**(proposer or PROPOSER_DEFAULT))
        cfg = ServerCfg(hbm_budget, max_batch, int(learned), search_range, lambda8, pc)
        self._params = None
        if learned:
            self._params = np.ascontiguousarray(load_proposer_weights() if params is None else params, np.float32)
        self.h = hip.bc250_server_create(C.byref(cfg), _ptr(self._params) if learned else None)
        if not self.h:
            raise RuntimeError("bc250_server_create failed")

    def session_bytes(self, **kw):
        return hip.bc250_server_session_bytes(self.h, C.byref(SessionCfg(**kw)))

    def open(self, width, height, codec=GPU_CODEC_HEVC, qp=27, num_slices=1, gop=0, ring=1):
        """Session id, or -1 when its footprint does not fit the budget."""
        return hip.bc250_server_open(self.h, C.byref(SessionCfg(width, height, codec, qp, num_slices, gop, ring)))

    def close_session(self, sid):
        if hip.bc250_server_close(self.h, sid) != 0:
            raise RuntimeError("close failed")

    def upload(self, sid, source, y, uv):
        y, uv = np.ascontiguousarray(y), np.ascontiguousarray(uv)
        if hip.bc250_server_upload(self.h, sid, source, _ptr(y), y.strides[0], _ptr(uv), uv.strides[0]) != 0:
            raise RuntimeError("upload failed")

    def run(self, requests, checksums=False):
        """requests: (session, frame, source, t_arrival) sorted by arrival;
        returns (results as a structured numpy array, wall seconds)."""
        n = len(requests)
        reqs = (Request * n)(*[Request(s, f, src, 0, t) for s, f, src, t in requests])
        out = (FrameResult * n)()
        wall = C.c_double()
        if hip.bc250_server_run(self.h, reqs, n, int(checksums), out, C.byref(wall)) != 0:
            raise RuntimeError("bc250_server_run failed")
        dt = np.dtype([("session", "<i4"), ("frame", "<i4"), ("batch", "<i4"), ("fallbacks", "<i4"),
                       ("t_arrival", "<f8"), ("t_start", "<f8"), ("t_done", "<f8"), ("checksum", "<u8")])
        return np.frombuffer(bytes(out), dt).copy(), wall.value

    def autotune(self):
        if hip.bc250_server_autotune(self.h) != 0:
            raise RuntimeError("autotune failed")

    def stats(self):
        st = ServerStats()
        hip.bc250_server_stats(self.h, C.byref(st))
        return {"batches": st.batches, "frames": st.frames, "max_batch": st.max_batch_seen, "sessions": st.sessions,
                "mean_batch": st.mean_batch, "forward_variant": st.forward_variant,
                "autotune_us": list(st.autotune_us)[:3], "hbm_used": st.hbm_used}

    def close(self):
        if self.h:
            hip.bc250_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def staging_checksum(results, codec):
    """The server's FNV-1a over a frame's staging, computed from Encoder
    results (the same arrays in the same order)."""
    h = 0xcbf29ce484222325
    keys = ("quant", "mv8", "hevc_mode") if codec == GPU_CODEC_HEVC else ("quant", "mv8", "dc", "pred_mode", "mv")
    for k in keys:
        a = np.ascontiguousarray(results[k]).view(np.uint8)
        h = hip.bc250_fnv1a64(h, _ptr(a), a.size)
    return h
_sig(hip, "gpu_compute_get_me_fallbacks", _I, [_CTX, _I])
# the default search of the learned form: 4 candidates, fallback above
# 18/16 of the neighbours' median SAD + 128 or above 3000 (chosen on the
# 1080p bench sequence: 0.57% BD-rate against the plain search with 3.3% of
# P-picture MBs searched in full)
PROPOSER_DEFAULT = dict(ncand=4, alpha_q4=18, beta=128, t_high=3000)
WEIGHTS = os.path.join(_HERE, "weights", "me_proposer.f32")


def load_proposer_weights(path=WEIGHTS):
    """The trained proposer's parameters (raw little-endian float32)."""
    w = np.fromfile(path, dtype="<f4")
    if w.size != hip.bc250_cnn_param_count():
        raise ValueError(f"{path}: {w.size} parameters, expected {hip.bc250_cnn_param_count()}")
    return w
_sig(hip, "bc250_launch_me_refine", _I, [_P, _I, _P, _I, _I, _I, _P, _P, _P, _P, _I, _P])
_sig(hip, "bc250_launch_p_encode", _I, [_P, _P, _I, _I, _P, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P,
                                       _P, _P, _P])
_sig(hip, "bc250_launch_i_encode", _I, [_P, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _U,
                                       _P])
_sig(hip, "bc250_launch_deblock", _I, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _I, _I, _P, _U, _P])
_sig(hip, "bc250_launch_zero_i32", _I, [_P, _I, _P])
_sig(hip, "bc250_deblock_work_bytes", C.c_size_t, [_I, _I])
_sig(hip, "bc250_intra_work_bytes", C.c_size_t, [_I, _I])
_sig(hip, "bc250_launch_pad_edges", _I, [_P, _P, _I, _I, _I, _I, _I, _I, _P])
_sig(hip, "bc250_launch_square_me", _I, [_P, _I, _P, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_launch_diamond_me", _I, [_P, _I, _P, _I, _I, _I, _I, _P, _P])
_sig(hip, "cpu_simd_sad_16x16", _U, [_P, _I, _P, _I])
_sig(hip, "cpu_simd_me_search_frame", _I, [_P, _I, _P, _I, _U, _U, _P, _P])
_sig(hip, "bc250_h264_pred4x4_batch", _I, [_P, _P, _I, _P, _P])
_sig(hip, "bc250_h264_pred8x8_batch", _I, [_P, _P, _I, _P, _P])
_sig(hip, "bc250_h264_tq8x8_batch", _I, [_P, _I, _I, _I, _P, _P])
_sig(hip, "bc250_h264_dqit8x8_batch", _I, [_P, _I, _I, _P, _P])
# HEVC stages (include/bc250_hevc.h)
_sig(hip, "bc250_hevc_tq4_batch", _I, [_P, _I, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_dqit4_batch", _I, [_P, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_tq8_batch", _I, [_P, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_dqit8_batch", _I, [_P, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_tqn_batch", _I, [_P, _I, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_dqitn_batch", _I, [_P, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_tq32_batch", _I, [_P, _I, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_dqit32_batch", _I, [_P, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_mc_uni_batch", _I, [_P, _I, _I, _I, _P, _I, _I, _I, _I, _I, _P, _P])
_sig(hip, "bc250_hevc_launch_refine8", _I, [_P, _I, _P, _I, _I, _I, _P, _P, _I, _P])
_sig(hip, "bc250_hevc_intra_work_bytes", C.c_size_t, [_I, _I])
_sig(hip, "bc250_hevc_launch_i_encode", _I, [_P, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _U,
                                            _I, _P])
_sig(hip, "bc250_hevc_launch_p_encode", _I, [_P, _P, _I, _I, _P, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _P,
                                            _P, _I, _P])
_sig(hip, "bc250_hevc_launch_deblock", _I, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _I, _I, _P])
_sig(hip, "bc250_hevc_launch_deblock_ctu", _I, [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _I, _I, _I, _P])
_sig(hip, "bc250_hevc_cu_cbf_ctu", _I, [_P, _I, _I, _P, _P, _P])
_sig(hip, "bc250_hevc_launch_p_encode_ctu", _I, [_P, _P, _I, _I, _P, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _P,
                                                 _P, _I, _P, _P, _P, _I, _P])
_sig(hip, "bc250_hevc_launch_refine_p_encode", _I, [_P, _P, _I, _I, _P, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I,
                                                    _P, _P, _P, _I, _P, _P, _P, _I, _P])
_sig(hip, "bc250_hevc_launch_i_encode_ctu", _I, [_P, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P,
                                                 _U, _I, _I, _P])
_sig(hip, "bc250_hevc_cu_cbf", _I, [_P, _I, _P, _P])
_sig(hip, "bc250_hevc_p010_to_8bit", _I, [_P, _I, _P, _I, _I, _I, _P])

# H.264 CAVLC slices (include/bc250_bitstream.h)
_sig(hip, "bc250_h264_cavlc_work_size", C.c_size_t, [_I, _I])
_sig(hip, "bc250_h264_cavlc_slices", _I, [_P, _P, _P, _P, _P, _P, C.c_size_t, _P, _P, C.c_size_t, _P])
_sig(hip, "bc250_h264_cavlc_copy_out", _I, [_P, _P, _I, _P, C.c_size_t, _P, _P])
_sig(hip, "gpu_compute_set_h264_slice_header", _I, [_CTX, _P])
_sig(hip, "gpu_compute_get_slices_slot", _I, [_CTX, _I, C.POINTER(_P), C.POINTER(C.POINTER(C.c_uint32)),
                                              C.POINTER(_I)])

_sig(host, "bc250_gen_create", _P, [C.c_uint64, _I, _I, _I])
_sig(host, "bc250_gen_destroy", None, [_P])
_sig(host, "bc250_gen_frame", _I, [_P, _I, _P, _I, _P, _I])
_sig(host, "bc250_gen_offset_qpel", None, [_P, _I, C.POINTER(_I), C.POINTER(_I)])
_sig(host, "bc250_encode_sequence", _I, [_CTX, C.POINTER(GpuImage), _I, _I, _I, C.POINTER(SeqCfg),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)])
# per-frame consumer of bc250_encode_sequence_cb: (user, frame, slot, is_intra) -> int
FRAME_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int)
_sig(host, "bc250_encode_sequence_cb", _I, [_CTX, C.POINTER(GpuImage), _I, _I, _I, C.POINTER(SeqCfg),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p, C.c_void_p])


class NalSink(C.Structure):
    """bc250_nal_sink_t (host/bc250_host.h)."""
    _fields_ = [("ctx", C.c_void_p), ("buf", C.c_void_p), ("cap", C.c_size_t), ("used", C.c_size_t),
                ("frame_bytes", C.c_void_p)]


NAL_SINK_FRAME = C.cast(host.bc250_nal_sink_frame, C.c_void_p)
_sig(host, "bc250_encode_sequence_rc", _I, [_CTX, C.POINTER(GpuImage), _I, _I, _I, C.POINTER(SeqCfg), _P, _P, _P,
                                           _P, C.POINTER(C.c_double), C.POINTER(C.c_double)])
_sig(host, "h264_intra16_luma_dc_transform", None, [_P, _I, _P, _P])
_sig(host, "h264_chroma_dc_levels", None, [_P, _I, _P])
_sig(host, "h264_frame_dc_levels", None, [_P, _I, _I, _I, _P, _P])
_sig(host, "h264_chroma_qp", _I, [_I])
_sig(host, "bc250_rbsp_to_ebsp", C.c_size_t, [_P, C.c_size_t, _P, C.c_size_t])
_sig(host, "bc250_h264_frame_nals", C.c_size_t, [_P, _P, _I, _I, _P, C.c_size_t])
_sig(host, "bc250_h264_write_sps", C.c_size_t, [_P, _P, C.c_size_t])
_sig(host, "bc250_h264_write_pps", C.c_size_t, [_P, _P, C.c_size_t])


class HevcStreamCfg(C.Structure):
    """bc250_hevc_stream_cfg_t (include/bc250_hevc_stream.h)."""
    _fields_ = [(k, C.c_int32) for k in ("width", "height", "crop_width", "crop_height", "bit_depth", "ctu_log2",
                                         "num_slices", "deblock", "init_qp")]


_sig(host, "bc250_hevc_write_parameter_sets", C.c_size_t, [C.POINTER(HevcStreamCfg), _P, C.c_size_t])
_sig(host, "bc250_hevc_write_picture", C.c_size_t, [C.POINTER(HevcStreamCfg), _I, _I, _I, _P, _P, _P, _P,
                                                    C.c_size_t])


class HevcStream:
    """The HEVC bitstream of pipeline pictures (include/bc250_hevc_stream.h):
    headers() gives VPS / SPS / PPS, picture() one picture's slice NAL units
    from its staging (levels int16, modes uint8, 8x8 CU vectors uint32 x 4),
    both in Annex B form."""

    def __init__(self, width, height, bit_depth=8, ctu_log2=4, num_slices=1, deblock=True, init_qp=26,
                 crop_width=0, crop_height=0):
        self.cfg = HevcStreamCfg(width, height, crop_width, crop_height, bit_depth, ctu_log2, num_slices,
                                 int(bool(deblock)), init_qp)
        self.w, self.h = width, height

    def headers(self):
        buf = np.empty(1024, np.uint8)
        n = host.bc250_hevc_write_parameter_sets(C.byref(self.cfg), _ptr(buf), buf.size)
        if not n:
            raise ValueError("invalid HEVC stream configuration")
        return buf[:n].tobytes()

    def picture(self, is_idr, poc, qp, coef, modes, mv8=None):
        coef = np.ascontiguousarray(coef, np.int16)
        modes = np.ascontiguousarray(modes, np.uint8)
        t = (self.w // 16) * (self.h // 16)
        if coef.size != t * 384 or modes.size != t * 16:
            raise ValueError("staging does not match the picture size")
        mvp = None
        if not is_idr:
            mv = np.ascontiguousarray(mv8).view(np.uint32).reshape(-1)
            if mv.size != t * 16:
                raise ValueError("vectors do not match the picture size")
            mvp = _ptr(mv)
        cap = self.w * self.h * 8 + (1 << 16)
        buf = np.empty(cap, np.uint8)
        n = host.bc250_hevc_write_picture(C.byref(self.cfg), int(bool(is_idr)), poc, qp, _ptr(coef), _ptr(modes), mvp,
                                          _ptr(buf), cap)
        if not n:
            raise RuntimeError("HEVC picture writer failed")
        return buf[:n].tobytes()


class RcState(C.Structure):
    """bc250_rc_t (host/bc250_rc.h)."""
    _fields_ = [("mode", C.c_int32), ("target_bitrate", C.c_uint32), ("max_bitrate", C.c_uint32),
                ("qp_min", C.c_int32), ("qp_max", C.c_int32), ("framerate", C.c_double),
                ("target_bits_per_frame", C.c_uint32), ("buffer_fullness", C.c_int64), ("buffer_size", C.c_int64),
                ("base_qp", C.c_int32), ("current_qp", C.c_int32), ("prev_frame_sad", C.c_uint64),
                ("error_integral", C.c_int64), ("max_frame_bits", C.c_uint32), ("quality_level", C.c_uint32),
                ("wallclock_drain", C.c_int32), ("last_frame_ns", C.c_uint64), ("measured_fps", C.c_double),
                ("model", C.c_int32), ("cplx", C.c_double * 2), ("debt", C.c_double), ("model_last_p_qp", C.c_int32),
                ("cplx_still", C.c_double), ("still_last_qp", C.c_int32), ("still_share", C.c_double),
                ("pixels", C.c_double)]


RC_CBR, RC_VBR, RC_LOW_LATENCY, RC_CQP = 0, 1, 2, 3
_sig(host, "bc250_rc_init", None, [_P, _I, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32])
_sig(host, "bc250_rc_update_bitrate", None, [_P, C.c_uint32, C.c_uint32, C.c_uint32])
_sig(host, "bc250_rc_get_frame_qp", _I, [_P, C.c_uint64])
_sig(host, "bc250_rc_update_stats", None, [_P, _I])
_sig(host, "bc250_rc_model_frame_qp", _I, [_P, _I, _I])
_sig(host, "bc250_rc_model_frame_coded", None, [_P, _I, _I, _I, _I])
_sig(host, "bc250_rc_set_max_frame_size", None, [_P, C.c_uint32])
_sig(host, "bc250_rc_set_quality_level", None, [_P, C.c_uint32])
_sig(host, "bc250_rc_set_wallclock_drain", None, [_P, _I])


class RateControl:
    """The host rate controller (host/bc250_rc.c): the reference's
    rate_control.c interface, method for function."""

    def __init__(self, mode, bitrate, fps, width, height, model=False, current_qp=0):
        self.st = RcState()
        self.st.current_qp = current_qp   # the QP a CQP controller keeps
        host.bc250_rc_init(C.byref(self.st), mode, bitrate, fps, width, height)
        self.st.model = int(bool(model))

    def frame_qp(self, est_sad=0, intra=False, still=False):
        if self.st.model:
            return host.bc250_rc_model_frame_qp(C.byref(self.st), int(intra), int(still))
        return host.bc250_rc_get_frame_qp(C.byref(self.st), est_sad)

    def coded(self, bits, qp=0, intra=False, still=False):
        host.bc250_rc_update_stats(C.byref(self.st), int(bits))
        if self.st.model:
            host.bc250_rc_model_frame_coded(C.byref(self.st), int(intra), int(still), qp, int(bits))

    def update_bitrate(self, bitrate, width, height):
        host.bc250_rc_update_bitrate(C.byref(self.st), bitrate, width, height)

    def set_max_frame_size(self, bits):
        host.bc250_rc_set_max_frame_size(C.byref(self.st), bits)

    def set_quality_level(self, level):
        host.bc250_rc_set_quality_level(C.byref(self.st), level)

# bc250_h264_sps_t / bc250_h264_pps_t, field by field (all int32)
SPS_FIELDS = ("profile_idc", "level_idc", "sps_id", "chroma_format_idc", "bit_depth_luma", "bit_depth_chroma",
              "log2_max_frame_num_minus4", "pic_order_cnt_type", "log2_max_poc_lsb_minus4", "max_num_ref_frames",
              "width_in_mbs", "height_in_mbs", "frame_mbs_only", "direct_8x8_inference", "frame_cropping",
              "crop_left", "crop_right", "crop_top", "crop_bottom", "vui_present", "sar_width", "sar_height",
              "timing_info_present", "num_units_in_tick", "time_scale")
PPS_FIELDS = ("pps_id", "sps_id", "entropy_coding_mode", "pic_order_present", "num_ref_idx_l0_default",
              "num_ref_idx_l1_default", "weighted_pred", "weighted_bipred_idc", "pic_init_qp_minus26",
              "chroma_qp_offset", "second_chroma_qp_offset", "deblocking_filter_control", "constrained_intra_pred",
              "transform_8x8_mode")


def h264_parameter_sets(sps, pps):
    """SPS and PPS NAL units (Annex B) from field dicts (missing fields 0)."""
    out = b""
    for fields, vals, f in ((SPS_FIELDS, sps, host.bc250_h264_write_sps), (PPS_FIELDS, pps, host.bc250_h264_write_pps)):
        a = np.array([int(vals.get(k, 0)) for k in fields], np.int32)
        buf = np.zeros(512, np.uint8)
        n = f(_ptr(a), _ptr(buf), buf.size)
        if n == 0:
            raise RuntimeError("parameter set does not fit")
        out += buf[:n].tobytes()
    return out


def rbsp_to_ebsp(rbsp):
    """Emulation prevention of one RBSP (host/h264_host.c)."""
    src = np.frombuffer(bytes(rbsp), np.uint8)
    dst = np.zeros(len(src) * 3 // 2 + 2, np.uint8)
    n = host.bc250_rbsp_to_ebsp(_ptr(dst), dst.size, _ptr(src) if src.size else None, src.size)
    return dst[:n].tobytes()


def h264_frame_nals(slices, is_idr):
    """Annex B NAL units of a frame's slice RBSPs (list of bytes)."""
    data = b"".join(slices)
    info = np.zeros(4 * len(slices) + 2, np.uint32)
    off = 0
    for s, x in enumerate(slices):
        info[4 * s], info[4 * s + 1] = off, len(x)
        off += len(x)
    src = np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8)
    dst = np.zeros(2 * len(data) + 8 * len(slices) + 16, np.uint8)
    n = host.bc250_h264_frame_nals(_ptr(src), _ptr(info), len(slices), int(is_idr), _ptr(dst), dst.size)
    if n == 0:
        raise RuntimeError("NAL buffer too small")
    return dst[:n].tobytes()


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


_HIPRT = None


def _hiprt():
    """The HIP runtime (already loaded by libbc250_hip.so)."""
    global _HIPRT
    if _HIPRT is None:
        _HIPRT = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
        _HIPRT.hipMemcpy2D.restype = C.c_int
        _HIPRT.hipMemcpy2D.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t,
                                       C.c_int]
    return _HIPRT


# ---------------------------------------------------------------- CAVLC

# bc250_h264_slice_cfg_t, field by field (all int32)
SLICE_CFG_FIELDS = ("width_in_mbs", "height_in_mbs", "qp", "is_idr", "num_slices", "pps_id", "pic_init_qp_minus26",
                    "frame_num", "log2_max_frame_num_minus4", "idr_pic_id", "poc", "log2_max_poc_lsb_minus4",
                    "deblock_idc", "relax_chroma_skip")
MAX_SLICES = 64


def h264_slice_cfg(width_in_mbs, height_in_mbs, qp, is_idr, num_slices=1, **kw):
    """A bc250_h264_slice_cfg_t as an int32 array; other fields by keyword
    (default 0)."""
    v = dict(width_in_mbs=width_in_mbs, height_in_mbs=height_in_mbs, qp=qp, is_idr=int(is_idr),
             num_slices=num_slices)
    unknown = set(kw) - set(SLICE_CFG_FIELDS)
    if unknown:
        raise TypeError(f"unknown slice fields {sorted(unknown)}")
    v.update(kw)
    return np.array([int(v.get(k, 0)) for k in SLICE_CFG_FIELDS], np.int32)


class CavlcWriter:
    """The GPU H.264 CAVLC slice writer (bc250_h264_cavlc_slices) for frames
    of one size on the current device; owns its work and output buffers.
    Inputs are device tensors in the staging layout of include/gpu_compute.h:
    levels int16 [nmb*384], dc int32 [nmb*24], modes int32 [nmb], mvs int32
    [nmb*4] (gpu_mv_t)."""

    def __init__(self, width_in_mbs, height_in_mbs, cap_bytes=None):
        import torch
        self.torch = torch
        self.wmb, self.hmb = width_in_mbs, height_in_mbs
        nmb = width_in_mbs * height_in_mbs
        self.work = torch.empty(int(hip.bc250_h264_cavlc_work_size(width_in_mbs, height_in_mbs)),
                                dtype=torch.uint8, device="cuda")
        cap = cap_bytes if cap_bytes is not None else nmb * 2560 + 4096 * MAX_SLICES
        self.cap = (cap + 3) // 4 * 4
        self.out = torch.empty(self.cap // 4, dtype=torch.int32, device="cuda")
        self.info = torch.zeros(4 * MAX_SLICES + 2, dtype=torch.int32, device="cuda")

    def launch(self, levels, dc, modes, mvs, cfg, stream=None):
        """Enqueues the frame's slices on `stream` (a torch stream, or the
        current one); raises on invalid arguments."""
        cfg = np.ascontiguousarray(cfg, np.int32)
        if cfg[0] != self.wmb or cfg[1] != self.hmb:
            raise ValueError("cfg size differs from the writer's")
        st = (stream or self.torch.cuda.current_stream()).cuda_stream

        def ptr(t):
            return None if t is None else C.c_void_p(t.data_ptr())
        rc = hip.bc250_h264_cavlc_slices(ptr(levels), ptr(dc), ptr(modes), ptr(mvs), _ptr(cfg),
                                         ptr(self.out), self.cap, ptr(self.info), ptr(self.work),
                                         self.work.numel(), C.c_void_p(st))
        if rc != 0:
            raise ValueError("bc250_h264_cavlc_slices rejected its arguments")

    def result(self, num_slices):
        """Synchronizes and returns the slices' RBSPs (list of bytes)."""
        self.torch.cuda.synchronize()
        info = self.info.cpu().numpy().view(np.uint32)
        if info[4 * num_slices + 1]:
            raise RuntimeError(f"CAVLC output exceeded {self.cap} bytes")
        total = int(info[4 * num_slices])
        data = self.out.view(self.torch.uint8)[:total].cpu().numpy().tobytes()
        return [data[info[4 * s]:info[4 * s] + info[4 * s + 1]] for s in range(num_slices)]

    def slices(self, levels, dc, modes, mvs, cfg):
        self.launch(levels, dc, modes, mvs, cfg)
        return self.result(int(cfg[4]))


# ---------------------------------------------------------------- frames

class FrameGen:
    """Seeded synthetic NV12 frames (host/framegen.c)."""

    def __init__(self, seed, width, height, nrect):
        self.w, self.h = width, height
        self._g = host.bc250_gen_create(seed, width, height, nrect)
        if not self._g:
            raise ValueError("bad generator arguments")

    def frame(self, f):
        y = np.empty((self.h, self.w), np.uint8)
        uv = np.empty(((self.h + 1) // 2, 2 * ((self.w + 1) // 2)), np.uint8)
        host.bc250_gen_frame(self._g, f, _ptr(y), y.strides[0], _ptr(uv), uv.strides[0])
        return y, uv

    def offset_qpel(self, f):
        ox, oy = C.c_int(), C.c_int()
        host.bc250_gen_offset_qpel(self._g, f, C.byref(ox), C.byref(oy))
        return ox.value, oy.value

    def close(self):
        if self._g:
            host.bc250_gen_destroy(self._g)
            self._g = None

    def __del__(self):
        self.close()


def pad_coded(y, uv):
    """Visible NV12 planes (uint8, or uint16 10-bit samples) -> coded-size
    planes (16-aligned) with the last row/column replicated, which is what an
    upload leaves on the device."""
    h, w = y.shape
    wc, hc = (w + 15) // 16 * 16, (h + 15) // 16 * 16
    yc = np.pad(y, ((0, hc - h), (0, wc - w)), mode="edge")
    u = uv[:, 0::2]
    v = uv[:, 1::2]
    u = np.pad(u, ((0, hc // 2 - u.shape[0]), (0, wc // 2 - u.shape[1])), mode="edge")
    v = np.pad(v, ((0, hc // 2 - v.shape[0]), (0, wc // 2 - v.shape[1])), mode="edge")
    uvc = np.empty((hc // 2, wc), y.dtype)
    uvc[:, 0::2] = u
    uvc[:, 1::2] = v
    return np.ascontiguousarray(yc), uvc


# ---------------------------------------------------------------- encoder

class Encoder:
    """One gpu_compute context: upload frames, encode, read staging back."""

    def __init__(self, width, height, search_range=64, lambda8=1, deblock_idc=0, me_8x8=1, readback=1, codec=0,
                 bit_depth=8, square_radius=0, diamond=0, ctu_log2=4, entropy=0):
        """bit_depth 10: P010 surfaces (HEVC Main10); upload() then takes uint16
        planes of 0..1023 samples and map_recon() returns them.  ctu_log2 5:
        HEVC CTU32 (gpu_encode_params_t.hevc_ctu_log2).  entropy 1: H.264
        frames end with their CAVLC slices, read with slices()."""
        self.w, self.h = width, height
        self.fmt = GPU_IMAGE_P010 if bit_depth == 10 else GPU_IMAGE_NV12
        self.wc, self.hc = (width + 15) // 16 * 16, (height + 15) // 16 * 16
        self.nmb = (self.wc // 16) * (self.hc // 16)
        self.ctx = GpuContext()
        if hip.bc250_gpu_init(C.byref(self.ctx)) != 0:
            raise RuntimeError("bc250_gpu_init failed (no HIP device?)")
        self.codec = codec
        self.entropy = entropy
        self.set_params(search_range, lambda8, deblock_idc, me_8x8, readback, codec, square_radius, diamond, ctu_log2,
                        entropy)
        self.images = []

    def set_params(self, search_range=64, lambda8=1, deblock_idc=0, me_8x8=1, readback=1, codec=0, square_radius=0,
                   diamond=0, ctu_log2=4, entropy=0):
        """codec: GPU_CODEC_H264 (0) or GPU_CODEC_HEVC (1); square_radius 2..16
        selects the reference's Tier-2 square search, diamond 1 its GPU
        diamond search, instead of the full one."""
        p = EncodeParams(search_range, lambda8, deblock_idc, me_8x8, readback, codec, square_radius, diamond, ctu_log2,
                         entropy)
        if hip.gpu_compute_set_encode_params(C.byref(self.ctx), C.byref(p)) != 0:
            raise ValueError("invalid encode parameters")
        self.codec, self.entropy = codec, entropy

    def new_image(self):
        img, mem = GpuImage(), GpuMemory()
        if hip.gpu_compute_create_image(C.byref(self.ctx), self.w, self.h, self.fmt, C.byref(img),
                                        C.byref(mem)) != 0:
            raise RuntimeError("create_image failed")
        self.images.append((img, mem))
        return img, mem

    def upload(self, img_mem, y, uv):
        img, mem = img_mem
        if self.fmt == GPU_IMAGE_P010:
            y = np.ascontiguousarray(y, np.uint16)
            uv = np.ascontiguousarray(uv, np.uint16)
            f = hip.gpu_compute_upload_p010
        else:
            y = np.ascontiguousarray(y)
            uv = np.ascontiguousarray(uv)
            f = hip.gpu_compute_upload_nv12
        if f(C.byref(self.ctx), C.byref(img), mem, _ptr(y), y.strides[0], _ptr(uv), uv.strides[0], self.w,
             self.h) != 0:
            raise RuntimeError("upload failed")

    def encode(self, img_mem, qp, is_intra, num_slices=1, cpu_mvs=None, me_mode=0):
        """Submit one frame and wait for it; returns the slot."""
        img, _ = img_mem
        c = C.byref(self.ctx)
        if hip.gpu_compute_begin_picture(c, img) != 0:
            raise RuntimeError("begin_picture failed")
        mvp = None
        if cpu_mvs is not None:
            cpu_mvs = np.ascontiguousarray(cpu_mvs, dtype=MV_DTYPE)
            mvp = _ptr(cpu_mvs)
        if hip.gpu_compute_dispatch_encode_ext(c, img, self.w, self.h, qp, int(is_intra), num_slices, me_mode,
                                               mvp) != 0:
            raise RuntimeError("dispatch_encode_ext failed")
        if hip.gpu_compute_end_picture(c) != 0:
            raise RuntimeError("end_picture failed")
        slot = hip.gpu_compute_submitted_slot(c)
        if hip.gpu_compute_sync_slot(c, slot) != 0:
            raise RuntimeError("sync_slot failed (error flags %d)" % hip.gpu_compute_take_error_flags(c))
        return slot

    def staging(self, slot, name):
        d, n = C.c_void_p(), C.c_size_t()
        f = getattr(hip, f"gpu_compute_get_{name}_staging_data_slot")
        if f(C.byref(self.ctx), slot, C.byref(d), C.byref(n)) != 0:
            raise RuntimeError(f"no {name} staging data in slot {slot}")
        dt = {"quant": np.int16, "dc": np.int32, "pred_mode": np.uint32, "nz": np.uint32, "mv": MV_DTYPE,
              "mv8": MV_DTYPE, "hevc_mode": np.uint8}[name]
        buf = (C.c_uint8 * n.value).from_address(d.value)
        return np.frombuffer(buf, dtype=dt).copy()

    def slices(self, slot):
        """The CAVLC slices (list of bytes) of the frame in `slot` (entropy 1)."""
        d, info, n = C.c_void_p(), C.POINTER(C.c_uint32)(), C.c_int()
        if hip.gpu_compute_get_slices_slot(C.byref(self.ctx), slot, C.byref(d), C.byref(info), C.byref(n)) != 0:
            raise RuntimeError(f"no slices in slot {slot}")
        ns = n.value
        total = info[4 * ns]
        data = C.string_at(d.value, total)
        return [data[info[4 * s]:info[4 * s] + info[4 * s + 1]] for s in range(ns)]

    def set_proposer(self, params, ncand=4, alpha_q4=18, beta=128, t_high=3000):
        """The learned integer search (gpu_compute_set_me_proposer); params
        None turns it off."""
        if params is None:
            rc = hip.gpu_compute_set_me_proposer(C.byref(self.ctx), None, None)
        else:
            self._prop = np.ascontiguousarray(params, np.float32)
            cfg = ProposerCfg(ncand, alpha_q4, beta, t_high)
            rc = hip.gpu_compute_set_me_proposer(C.byref(self.ctx), _ptr(self._prop), C.byref(cfg))
        if rc != 0:
            raise ValueError("invalid proposer settings")

    def fallbacks(self, slot):
        """MBs the frame in `slot` searched in full after the learned search
        (-1: the frame did not use it)."""
        return hip.gpu_compute_get_me_fallbacks(C.byref(self.ctx), slot)

    def recon_luma_device(self, torch):
        """A device copy (torch uint8 [coded h, coded w]) of the newest
        reconstruction's luma (8-bit pictures)."""
        im = self.ctx.recon_image
        out = torch.empty((im.coded_height, im.coded_width), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        if _hiprt().hipMemcpy2D(C.c_void_p(out.data_ptr()), C.c_size_t(im.coded_width), C.c_void_p(im.y_plane),
                                C.c_size_t(im.y_pitch), C.c_size_t(im.coded_width), C.c_size_t(im.coded_height),
                                3) != 0:
            raise RuntimeError("hipMemcpy2D failed")
        return out

    def set_slice_header(self, **kw):
        """gpu_compute_set_h264_slice_header: pps_id, pic_init_qp_minus26,
        log2_max_frame_num_minus4, log2_max_poc_lsb_minus4, relax_chroma_skip."""
        cfg = h264_slice_cfg(1, 1, 26, 0, 1, **kw)
        if hip.gpu_compute_set_h264_slice_header(C.byref(self.ctx), _ptr(cfg)) != 0:
            raise ValueError("invalid slice header fields")

    def results(self, slot):
        if self.codec == GPU_CODEC_HEVC:
            return {k: self.staging(slot, k) for k in ("quant", "mv8", "hevc_mode")}
        return {k: self.staging(slot, k) for k in ("quant", "dc", "pred_mode", "mv", "nz", "mv8")}

    def recon(self):
        y = np.empty((self.h, self.w), np.uint8)
        uv = np.empty(((self.h + 1) // 2, 2 * ((self.w + 1) // 2)), np.uint8)
        if hip.gpu_compute_download_recon(C.byref(self.ctx), _ptr(y), y.strides[0], _ptr(uv), uv.strides[0]) != 0:
            raise RuntimeError("download_recon failed")
        return y, uv

    def map_recon(self):
        """Reads the newest reconstruction the way the reference's CPU ME route
        does (encoder_h264.c:2325-2378): cpu_read_safe, then map_surface on
        ctx.recon_image / ctx.recon_memory, then unmap.  Returns coded-size
        (y, uv) copies (P010: uint16 0..1023 samples)."""
        import numpy as np
        c = C.byref(self.ctx)
        if hip.gpu_compute_cpu_read_safe(c) != 1:
            raise RuntimeError("GPU work still outstanding")
        mem = self.ctx.recon_memory
        if hip.gpu_compute_wait_for_image_ready_host(c, mem) != 0:
            raise RuntimeError("recon surface not ready")
        lay, unmap = Nv12Layout(), C.c_bool(False)
        base = hip.gpu_compute_map_surface(c, C.byref(self.ctx.recon_image), mem, C.byref(lay), C.byref(unmap))
        if not base:
            raise RuntimeError("map_surface failed")
        try:
            im = self.ctx.recon_image
            bps = 2 if im.format == GPU_IMAGE_P010 else 1
            raw = np.ctypeslib.as_array((C.c_uint8 * lay.total_size).from_address(base))
            y = raw[lay.y_offset:lay.y_offset + lay.y_pitch * im.coded_height].reshape(
                im.coded_height, lay.y_pitch)[:, :im.coded_width * bps].copy()
            uv = raw[lay.uv_offset:lay.uv_offset + lay.uv_pitch * (im.coded_height // 2)].reshape(
                im.coded_height // 2, lay.uv_pitch)[:, :im.coded_width * bps].copy()
        finally:
            hip.gpu_compute_unmap_surface(c, mem, unmap)
        if bps == 2:   # P010 words -> 0..1023 samples
            y, uv = y.view(np.uint16), uv.view(np.uint16)
            if (y & 63).any() or (uv & 63).any():
                raise RuntimeError("P010 reconstruction has bits below the sample")
            y, uv = y >> 6, uv >> 6
        return y, uv

    def set_stage_timing(self, on=True):
        """Stage events on the context's stream (stage_times / stage_stats)."""
        if hip.gpu_compute_set_stage_timing(C.byref(self.ctx), int(bool(on))) != 0:
            raise RuntimeError("set_stage_timing failed")

    def stage_times(self, slot):
        ms = (C.c_float * 6)()
        if hip.gpu_compute_get_stage_times(C.byref(self.ctx), slot, ms) != 0:
            raise RuntimeError("stage times unavailable")
        return list(ms)

    def stage_stats(self, reset=False):
        st = StageStats()
        if hip.gpu_compute_get_stage_stats(C.byref(self.ctx), C.byref(st), int(reset)) != 0:
            raise RuntimeError("stage stats unavailable")
        return {"sum_ms": list(st.sum_ms), "total_ms": st.total_ms, "frames": st.frames, "p_frames": st.p_frames,
                "me_int_ms_p": st.me_int_ms_p}

    def error_flags(self):
        return hip.gpu_compute_take_error_flags(C.byref(self.ctx))

    def encode_sequence(self, images, qp, gop=0, num_slices=1, pipelined=1, collect=False, nal_sink=None):
        """Runs the host sequence driver (bc250_encode_sequence_cb) over
        `images`; returns (frame_ms, total_ms), plus, with collect, the list
        of every frame's outputs read in the driver's per-frame consumer
        (results() and, with entropy 1, "slices" and "is_intra").  nal_sink (a
        NalSink) makes the C NAL writer the consumer instead."""
        n = len(images)
        arr = (GpuImage * n)(*[im for im, _ in images])
        cfg = SeqCfg(qp, gop, num_slices, pipelined)
        fms = (C.c_double * n)()
        tot = C.c_double()
        got = [None] * n
        cb, user = None, None
        if collect:
            def on_frame(_user, frame, slot, is_intra):
                try:
                    r = self.results(slot)
                    if self.entropy:
                        r["slices"] = self.slices(slot)
                    r["is_intra"] = bool(is_intra)
                    got[frame] = r
                    return 0
                except Exception as e:  # reported by the driver's return code
                    got[frame] = e
                    return 1
            cb = FRAME_CB(on_frame)
        elif nal_sink is not None:
            nal_sink.ctx = C.addressof(self.ctx)
            cb, user = NAL_SINK_FRAME, C.c_void_p(C.addressof(nal_sink))
        cbp = C.cast(cb, C.c_void_p) if cb is not None else None
        rc = host.bc250_encode_sequence_cb(C.byref(self.ctx), arr, n, self.w, self.h, C.byref(cfg), fms, C.byref(tot),
                                          cbp, user)
        if rc != 0:
            bad = [g for g in got if isinstance(g, Exception)]
            raise RuntimeError("encode_sequence failed" + (f": {bad[0]}" if bad else ""))
        out = np.array(fms[:], dtype=np.float64), tot.value
        return out + (got,) if collect else out

    def encode_sequence_rc(self, images, rc, gop=0, num_slices=1, pipelined=1, cap=None):
        """Rate-controlled H.264 sequence (bc250_encode_sequence_rc; the
        encoder needs entropy 1): rc is a RateControl.  Returns (stream bytes,
        per-frame QPs, per-frame bits, frame_ms, total_ms)."""
        if not self.entropy:
            raise ValueError("rate control needs the CAVLC slices (entropy=1)")
        n = len(images)
        arr = (GpuImage * n)(*[im for im, _ in images])
        cfg = SeqCfg(0, gop, num_slices, pipelined)
        buf = np.zeros(cap or (self.w * self.h * 2 * n + 4096), np.uint8)
        fbytes = np.zeros(n, np.uint64)
        sink = NalSink(C.addressof(self.ctx), buf.ctypes.data, buf.size, 0, fbytes.ctypes.data)
        qps, bits = np.zeros(n, np.int32), np.zeros(n, np.int32)
        fms = (C.c_double * n)()
        tot = C.c_double()
        r = host.bc250_encode_sequence_rc(C.byref(self.ctx), arr, n, self.w, self.h, C.byref(cfg), C.byref(rc.st),
                                          C.byref(sink), _ptr(qps), _ptr(bits), fms, C.byref(tot))
        if r != 0:
            raise RuntimeError("encode_sequence_rc failed")
        return buf[:sink.used].tobytes(), qps, bits, np.array(fms[:], np.float64), tot.value

    def close(self):
        if self.ctx.hip:
            for img, mem in self.images:
                hip.gpu_compute_destroy_image(C.byref(self.ctx), img, mem)
            self.images = []
            hip.bc250_gpu_destroy(C.byref(self.ctx))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
