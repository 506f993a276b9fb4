"""The learned motion proposer from Python: device buffers (torch) and the
gfx950 kernels of csrc/me_cnn.hip / csrc/me_learned.hip through the C ABI
(include/bc250_learned.h), for training and for the tests.  The encoder's
own inference path is native (gpu_compute_set_me_proposer); this module
trains the weights it loads.

Layer l (1..4) of the network: conv 3x3 (2->16 and 16->32 at stride 2,
32->32 twice at stride 1) + ReLU on 64x64 tiles, so conv4 is 16x16; then a
linear head per 16x16 block of the picture over the mean of its 4x4 cell of
conv4 and 16 neighbour vectors -> 4 hypotheses and their confidences."""
import ctypes as C

import numpy as np

LAYERS = ((2, 16), (16, 32), (32, 32), (32, 32))
NOUT, NFEAT = 12, 64
TP = 66                      # padded side of the input tile
SIDES = (66, 34, 18, 18, 18)   # padded side of activation 0 (input) .. 4


def kpad(cin):
    return (9 * cin + 31) // 32 * 32


def _off():
    off, o = [], 0
    for cin, cout in LAYERS:
        off.append(o)
        o += cout * (kpad(cin) + 1)
    off.append(o)
    return off, o + NOUT * (NFEAT + 1)


OFFSETS, NPARAM = _off()
PACKED_F = (0, 512, 10240, 28672)
PACKED_FLIP = (None, 5632, 19456, 37888)
NPACKED = 47104


def init_params(seed):
    """He-style initial weights, zero biases and padding columns (fp32 numpy)."""
    rs = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for l, (cin, cout) in enumerate(LAYERS):
        kp = kpad(cin)
        m = p[OFFSETS[l]:OFFSETS[l] + cout * (kp + 1)].reshape(cout, kp + 1)
        m[:, :9 * cin] = rs.normal(0, np.sqrt(2.0 / (9 * cin)), (cout, 9 * cin))
    h = p[OFFSETS[4]:].reshape(NOUT, NFEAT + 1)
    h[:, :NFEAT] = rs.normal(0, 0.05, (NOUT, NFEAT))
    return p


class Net:
    """Device state of the network for up to `ntiles` tiles per call."""

    def __init__(self, pkg, ntiles, params=None, train=False):
        import torch
        self.torch, self.hip, self.n = torch, pkg.hip, ntiles
        dev = "cuda"
        p = init_params(0) if params is None else np.asarray(params, np.float32)
        if p.size != NPARAM:
            raise ValueError("parameter vector has the wrong size")
        self.params = torch.from_numpy(p.copy()).to(dev)
        self.packed = torch.zeros(NPACKED, dtype=torch.int16, device=dev)
        z16 = lambda l, c: torch.zeros((ntiles, SIDES[l], SIDES[l], c), dtype=torch.int16, device=dev)
        self.act = [z16(0, 2)] + [z16(l + 1, c) for l, (_, c) in enumerate(LAYERS)]
        self.rows = torch.zeros((ntiles, 16, 4, 32), dtype=torch.float32, device=dev)
        self.pool = torch.zeros((ntiles, 16, 32), dtype=torch.float32, device=dev)
        self.z = torch.zeros((ntiles, 16, NOUT), dtype=torch.float32, device=dev)
        self.train = train
        if train:
            self.dz = [None] + [z16(l + 1, c) for l, (_, c) in enumerate(LAYERS)]
            self.grad = torch.zeros(NPARAM, dtype=torch.float32, device=dev)
            self.m = torch.zeros(NPARAM, dtype=torch.float32, device=dev)
            self.v = torch.zeros(NPARAM, dtype=torch.float32, device=dev)
            self.part = torch.zeros(ntiles * 32 * (kpad(32) + 1), dtype=torch.float32, device=dev)
            self.dpool = torch.zeros((ntiles, 16, 32), dtype=torch.float32, device=dev)
            self.loss = torch.zeros(ntiles, dtype=torch.float32, device=dev)
            self.step_no = 0
        self.pack()

    @staticmethod
    def _p(t):
        return None if t is None else C.c_void_p(t.data_ptr())

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed")

    def pack(self):
        self._ok(self.hip.bc250_cnn_pack(self._p(self.params), self._p(self.packed), None), "pack")

    def wf(self, l):  # forward bf16 weights of layer l (1..4)
        return C.c_void_p(self.packed.data_ptr() + 2 * PACKED_F[l - 1])

    def wflip(self, l):
        return C.c_void_p(self.packed.data_ptr() + 2 * PACKED_FLIP[l - 1])

    def bias(self, l):
        cin, _ = LAYERS[l - 1]
        return C.c_void_p(self.params.data_ptr() + 4 * (OFFSETS[l - 1] + kpad(cin)))

    def head_w(self):
        return C.c_void_p(self.params.data_ptr() + 4 * OFFSETS[4])

    def set_input(self, cur, ref, pitch, w, h, tiles, slot_stride=0):
        """cur / ref: device uint8 luma (tensors or pointers); tiles: int32
        [n, 3] (slot, tx, ty)."""
        t = self.torch.as_tensor(np.ascontiguousarray(tiles, np.int32).reshape(-1)).cuda()
        n = t.numel() // 3
        ptr = lambda a: C.c_void_p(a if isinstance(a, int) else a.data_ptr())
        self._ok(self.hip.bc250_cnn_input(ptr(cur), pitch, ptr(ref), pitch, slot_stride, w, h, self._p(t), n,
                                          self._p(self.act[0]), None), "input")
        return n

    def forward(self, n, mvin, keep=True):
        """Layers 1..4 on the first n tiles, cell sums and head outputs."""
        for l in (1, 2, 3):
            self._ok(self.hip.bc250_cnn_conv(l, 0, self._p(self.act[l - 1]), self.wf(l), self.bias(l),
                                             self._p(self.act[l]), None, None, n, None), f"conv{l}")
        self._ok(self.hip.bc250_cnn_conv(4, 2, self._p(self.act[3]), self.wf(4), self.bias(4),
                                         self._p(self.act[4]) if keep else None, None, self._p(self.rows), n, None),
                 "conv4")
        self._ok(self.hip.bc250_cnn_pool(self._p(self.rows), self._p(self.pool), n, None), "pool")
        self._ok(self.hip.bc250_cnn_head(self.head_w(), self._p(self.pool), self._p(mvin), self._p(self.z), n, None),
                 "head")
        return self.z[:n]

    def forward_fused(self, n, mvin):
        """Inference in one launch (layers 1..4 and the block sums on chip),
        then the head; the activations act[1..4] are not written."""
        self._ok(self.hip.bc250_cnn_forward(self._p(self.act[0]), self._p(self.packed), self._p(self.params),
                                            self._p(self.pool), n, None), "fused forward")
        self._ok(self.hip.bc250_cnn_head(self.head_w(), self._p(self.pool), self._p(mvin), self._p(self.z), n, None),
                 "head")
        return self.z[:n]

    def backward(self, n, mvin, target, scale):
        """Loss per tile and the parameter gradient (times scale) of the last
        forward(keep=True) on n tiles; target [n,16,2] float pel (device)."""
        hw = C.c_void_p(self.grad.data_ptr() + 4 * OFFSETS[4])
        self._ok(self.hip.bc250_cnn_head_train(self.head_w(), self._p(self.pool), self._p(mvin), self._p(target),
                                               self._p(self.loss), self._p(self.dpool), self._p(self.part), hw, n,
                                               float(scale), None), "head backward")
        self._ok(self.hip.bc250_cnn_pool_backward(self._p(self.dpool), self._p(self.act[4]), self._p(self.dz[4]), n,
                                                  None), "pool backward")
        for l in (4, 3, 2, 1):
            g = C.c_void_p(self.grad.data_ptr() + 4 * OFFSETS[l - 1])
            self._ok(self.hip.bc250_cnn_wgrad(l, self._p(self.act[l - 1]), self._p(self.dz[l]), self._p(self.part), g,
                                              n, None), f"wgrad{l}")
            if l > 1:
                self._ok(self.hip.bc250_cnn_conv(l, 1, self._p(self.dz[l]), self.wflip(l), None,
                                                 self._p(self.dz[l - 1]), self._p(self.act[l - 1]), None, n, None),
                         f"dgrad{l}")
        return self.loss[:n]

    def adam(self, lr):
        self.step_no += 1
        self._ok(self.hip.bc250_cnn_adam(self._p(self.params), self._p(self.grad), self._p(self.m), self._p(self.v),
                                         NPARAM, float(lr), self.step_no, None), "adam")
        self.pack()

    def params_numpy(self):
        return self.params.cpu().numpy()
