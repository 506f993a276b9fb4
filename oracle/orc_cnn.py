"""CPU oracle of the learned motion proposer's network (test infrastructure:
imported only by tests/, smoke() and bench.py's CPU leg).

A numpy restatement of the network, loss and gradients that the gfx950
kernels of csrc/me_cnn.hip compute, in the same parameter layout
(include/bc250_learned.h).  The reference has no such network: its nearest
structure is the HEVC encoder's start-candidate search
(hevc_enc_template.c:958-1043), so nothing here follows reference code; the
formulas are those of the convolution / ReLU / mean-pool / linear layers and
of the loss as defined in bc250_learned.h.

Two modes:
  * rounding="bf16": activations, weights and pre-activation gradients are
    rounded to bf16 at the points where the kernels store them, products
    accumulate in float64 -- the checker of the GPU results;
  * rounding=None: everything in float64, no rounding -- what the
    finite-difference test differentiates, to pin the backward formulas.

The backward is written as the adjoint of the forward im2col (scatter of
dz * w into the padded input), not as the GPU's convolution with flipped
weights, so a mistake in the flip or transpose fails against it."""
import numpy as np

NOUT, NFEAT, EPS = 12, 64, 0.02
LAYERS = ((2, 16), (16, 32), (32, 32), (32, 32))
STRIDES = (2, 2, 1, 1)   # conv1 64 -> 32, conv2 32 -> 16, conv3 / conv4 at 16x16
T = 64
CELL = 4                 # a 16x16 block of the picture is a 4x4 cell of conv4


def kpad(cin):
    return (9 * cin + 31) // 32 * 32


def offsets():
    off, o = [], 0
    for cin, cout in LAYERS:
        off.append(o)
        o += cout * (kpad(cin) + 1)
    off.append(o)
    o += NOUT * (NFEAT + 1)
    return off, o


OFFSETS, NPARAM = offsets()


def bf16(x):
    """Round float32 values to the nearest bf16 (ties to even), as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).reshape(a.shape)


def layer_params(p, l):
    """(W [cout, 3, 3, cin], b [cout]) of conv layer l (0..3)."""
    cin, cout = LAYERS[l]
    kp = kpad(cin)
    m = np.asarray(p[OFFSETS[l]:OFFSETS[l] + cout * (kp + 1)], np.float64).reshape(cout, kp + 1)
    return m[:, :9 * cin].reshape(cout, 3, 3, cin), m[:, kp]


def head_params(p):
    m = np.asarray(p[OFFSETS[4]:OFFSETS[4] + NOUT * (NFEAT + 1)], np.float64).reshape(NOUT, NFEAT + 1)
    return m[:, :NFEAT], m[:, NFEAT]


def init_params(seed):
    """He-style initialisation (zero biases, zero padding columns)."""
    rs = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for l, (cin, cout) in enumerate(LAYERS):
        kp = kpad(cin)
        m = p[OFFSETS[l]:OFFSETS[l] + cout * (kp + 1)].reshape(cout, kp + 1)
        m[:, :9 * cin] = rs.normal(0, np.sqrt(2.0 / (9 * cin)), (cout, 9 * cin))
    h = p[OFFSETS[4]:].reshape(NOUT, NFEAT + 1)
    h[:, :NFEAT] = rs.normal(0, 0.05, (NOUT, NFEAT))
    return p


def input_tiles(cur, ref, tiles):
    """(cur - 128) / 64 and (ref - 128) / 64 of 64x64 tiles (edge samples
    repeated past the picture), bf16; tiles = [(tx, ty), ...]."""
    h, w = cur.shape
    out = np.zeros((len(tiles), T, T, 2), np.float32)
    for i, (tx, ty) in enumerate(tiles):
        ys = np.minimum(np.arange(ty * T, ty * T + T), h - 1)
        xs = np.minimum(np.arange(tx * T, tx * T + T), w - 1)
        out[i, :, :, 0] = (cur[np.ix_(ys, xs)].astype(np.float32) - 128) / 64
        out[i, :, :, 1] = (ref[np.ix_(ys, xs)].astype(np.float32) - 128) / 64
    return bf16(out)


def _patches(x, stride=1):
    """x [n, s, s, c] -> [n, s/stride, s/stride, 3, 3, c]: the 3x3 window of
    every output pixel, output (y, x) reading input rows stride*y - 1 ..
    stride*y + 1 and columns likewise, zero outside."""
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    n, s, c = x.shape[0], x.shape[1], x.shape[3]
    o = s // stride
    out = np.empty((n, o, o, 3, 3, c), x.dtype)
    for ky in range(3):
        for kx in range(3):
            out[:, :, :, ky, kx, :] = xp[:, ky:ky + stride * o:stride, kx:kx + stride * o:stride, :]
    return out


def conv(x, w, b, stride=1):
    """3x3 convolution with a zero border: x [n,s,s,cin], w [cout,3,3,cin]
    -> [n,s/stride,s/stride,cout]."""
    pt = _patches(np.asarray(x, np.float64), stride)
    return np.einsum("nyxabc,oabc->nyxo", pt, w, optimize=True) + b


def conv_input_grad(dz, w, stride=1):
    """Adjoint of conv with respect to its input: scatter dz * w into the
    padded input (output pixel (y, x) onto the window it read) and crop."""
    n, o = dz.shape[0], dz.shape[1]
    s = o * stride
    cin = w.shape[3]
    dxp = np.zeros((n, s + 2, s + 2, cin))
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + stride * o:stride, kx:kx + stride * o:stride, :] += np.einsum(
                "nyxo,oc->nyxc", dz, w[:, ky, kx, :], optimize=True)
    return dxp[:, 1:s + 1, 1:s + 1, :]


def conv_weight_grad(x, dz, stride=1):
    pt = _patches(np.asarray(x, np.float64), stride)
    return np.einsum("nyxabc,nyxo->oabc", pt, dz, optimize=True), dz.sum(axis=(0, 1, 2))


def forward(p, x, mvin, rounding="bf16"):
    """Activations a0..a4 ([n,64,64,2], [n,32,32,16], [n,16,16,32] x 3), the
    4x4 cell sums of a4 [n,16,32] (cell j = 16x16 block j of the tile) and
    head outputs z [n,16,12] from the cell means."""
    r = bf16 if rounding == "bf16" else (lambda v: np.asarray(v, np.float64))
    acts = [r(x)]
    for l in range(4):
        w, b = layer_params(p, l)
        if rounding == "bf16":
            w = bf16(w.astype(np.float32)).astype(np.float64)
        acts.append(r(np.maximum(conv(acts[-1], w, b, STRIDES[l]), 0.0).astype(np.float32 if rounding else np.float64)))
    a4 = np.asarray(acts[4], np.float64)
    n = a4.shape[0]
    pool = a4.reshape(n, 4, CELL, 4, CELL, 32).sum(axis=(2, 4)).reshape(n, 16, 32)
    wh, bh = head_params(p)
    feat = np.concatenate([pool / CELL ** 2, np.asarray(mvin, np.float64)], axis=2)
    z = feat @ wh.T + bh
    return acts, pool, z


def _huber(d, delta=1.0 / 16):
    a = np.abs(d)
    val = np.where(a <= delta, 0.5 * d * d / delta, a - 0.5 * delta)
    grad = np.where(a <= delta, d / delta, np.sign(d))
    return val, grad


def loss_and_dz(z, target):
    """Per-block loss and d loss / d z (z [n,16,12], target [n,16,2] pel):
    the Huber loss (delta 1/16, in units of 16 pel) of the hypothesis nearest
    the target (first on ties), EPS times that of the others, plus the cross
    entropy of the confidence logits against the nearest one."""
    t = np.asarray(target, np.float64) / 16.0
    hx, gx = _huber(z[..., 0:8:2] - t[..., None, 0])
    hy, gy = _huber(z[..., 1:8:2] - t[..., None, 1])
    l = hx + hy
    best = np.argmin(l, axis=-1)
    wk = np.where(np.arange(4) == best[..., None], 1.0, EPS)
    logits = z[..., 8:12]
    mx = logits.max(axis=-1, keepdims=True)
    e = np.exp(logits - mx)
    se = e.sum(axis=-1, keepdims=True)
    logp = logits - mx - np.log(se)
    loss = (wk * l).sum(-1) - np.take_along_axis(logp, best[..., None], -1)[..., 0]
    dz = np.zeros_like(z)
    dz[..., 0:8:2] = wk * gx
    dz[..., 1:8:2] = wk * gy
    dz[..., 8:12] = e / se - (np.arange(4) == best[..., None])
    return loss, dz


def backward(p, acts, pool, mvin, dz_head, rounding="bf16"):
    """Gradient of the parameters (flat, like p) for d loss / d z = dz_head,
    and the stored pre-activation gradients dz1..dz4 (bf16-rounded with
    rounding="bf16", as the kernels store them)."""
    r = bf16 if rounding == "bf16" else (lambda v: np.asarray(v, np.float64))
    g = np.zeros(NPARAM, np.float64)
    wh, bh = head_params(p)
    n = pool.shape[0]
    feat = np.concatenate([pool / CELL ** 2, np.asarray(mvin, np.float64)], axis=2)
    gh = g[OFFSETS[4]:OFFSETS[4] + NOUT * (NFEAT + 1)].reshape(NOUT, NFEAT + 1)
    gh[:, :NFEAT] = np.einsum("nbo,nbi->oi", dz_head, feat)
    gh[:, NFEAT] = dz_head.sum(axis=(0, 1))
    dpool = (dz_head @ wh[:, :32]) / CELL ** 2                    # [n,16,32]
    da = np.repeat(np.repeat(dpool.reshape(n, 4, 4, 32), CELL, axis=1), CELL, axis=2)
    dzs = [None] * 5
    for l in range(3, -1, -1):
        a_out = np.asarray(acts[l + 1], np.float64)
        dz = np.where(a_out > 0, da, 0.0)
        if rounding == "bf16":
            dz = r(dz.astype(np.float32)).astype(np.float64)
        dzs[l + 1] = dz
        cin, cout = LAYERS[l]
        kp = kpad(cin)
        gw, gb = conv_weight_grad(acts[l], dz, STRIDES[l])
        gm = g[OFFSETS[l]:OFFSETS[l] + cout * (kp + 1)].reshape(cout, kp + 1)
        gm[:, :9 * cin] = gw.reshape(cout, 9 * cin)
        gm[:, kp] = gb
        if l > 0:
            w, _ = layer_params(p, l)
            if rounding == "bf16":
                w = bf16(w.astype(np.float32)).astype(np.float64)
            da = conv_input_grad(dz, w, STRIDES[l])
    return g, dzs


def total_loss(p, x, mvin, target):
    """Sum of the per-block losses in float64 (for finite differences)."""
    _, _, z = forward(p, x, mvin, rounding=None)
    return loss_and_dz(z, target)[0].sum()
