"""bc250_h264_cavlc_slices_noalloc (its work buffer a file-scope device
array, no allocation) against bc250_h264_cavlc_slices with a caller-owned
work buffer, through the C ABI: two calls in a row, the second on a smaller
frame, byte-equal slices and slice_info; frames above its limit refused."""
import ctypes as C

import numpy as np
import pytest

import orc
from conftest import have_gpu, load_pkg
from test_oracle_cavlc import synth_frame

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    if not have_gpu():
        pytest.skip("no GPU")
    p = load_pkg()
    p.hip.bc250_h264_cavlc_slices_noalloc.restype = C.c_int
    p.hip.bc250_h264_cavlc_slices_noalloc.argtypes = [C.c_void_p] * 6 + [C.c_size_t, C.c_void_p, C.c_void_p]
    p.hip.bc250_h264_cavlc_noalloc_max_mbs.restype = C.c_size_t
    return p


def _run(pkg, torch, wmb, hmb, ns, idr, seed, noalloc):
    rs = np.random.default_rng(seed)
    lv, dc, modes, mvs = synth_frame(rs, wmb, hmb, idr, big=True)
    n = wmb * hmb
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in
           (lv.reshape(-1), dc.reshape(-1), modes.view(np.int32), mvs.view(np.int32).reshape(-1))]
    cfg = orc.slice_cfg(wmb, hmb, 27, idr, ns, frame_num=3, poc=6, idr_pic_id=1)
    cap = n * 2560 + 4096 * 64
    out = torch.full((cap // 4,), -1, dtype=torch.int32, device="cuda")  # not zero: the writer clears what it uses
    info = torch.full((4 * 64 + 2,), -1, dtype=torch.int32, device="cuda")
    ptrs = [C.c_void_p(t.data_ptr()) for t in dev]
    cfgp = cfg.ctypes.data_as(C.c_void_p)
    if noalloc:
        rc = pkg.hip.bc250_h264_cavlc_slices_noalloc(*ptrs, cfgp, C.c_void_p(out.data_ptr()), cap,
                                                     C.c_void_p(info.data_ptr()), None)
    else:
        work = torch.empty(int(pkg.hip.bc250_h264_cavlc_work_size(wmb, hmb)), dtype=torch.uint8, device="cuda")
        rc = pkg.hip.bc250_h264_cavlc_slices(*ptrs, cfgp, C.c_void_p(out.data_ptr()), cap,
                                             C.c_void_p(info.data_ptr()), C.c_void_p(work.data_ptr()),
                                             work.numel(), None)
    assert rc == 0
    torch.cuda.synchronize()
    inf = info.cpu().numpy().view(np.uint32)[:4 * ns + 2]
    data = out.view(torch.uint8)[:int(inf[4 * ns])].cpu().numpy().tobytes()
    return inf, data


def test_noalloc_equals_the_work_buffer_entry_point(pkg):
    import torch
    # 1080p P frame, 4 slices, then a smaller IDR frame, 2 slices
    for wmb, hmb, ns, idr, seed in ((120, 68, 4, 0, 1), (20, 12, 2, 1, 2)):
        print(f"noalloc vs work buffer: {wmb}x{hmb} MBs, {ns} slices, idr {idr}", flush=True)
        a = _run(pkg, torch, wmb, hmb, ns, idr, seed, noalloc=False)
        b = _run(pkg, torch, wmb, hmb, ns, idr, seed, noalloc=True)
        np.testing.assert_array_equal(a[0], b[0])
        assert a[1] == b[1]
        assert len(a[1]) > 100


def test_noalloc_refuses_frames_above_its_limit(pkg):
    lim = int(pkg.hip.bc250_h264_cavlc_noalloc_max_mbs())
    assert lim >= 240 * 135  # 3840x2160 fits
    cfg = orc.slice_cfg(lim + 1, 1, 27, 1, 1)
    z = C.c_void_p(16)  # never dereferenced: the size check comes first
    assert pkg.hip.bc250_h264_cavlc_slices_noalloc(z, z, z, z, cfg.ctypes.data_as(C.c_void_p), z, 4096, z,
                                                   None) == -1
