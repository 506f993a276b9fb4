"""The learned proposer's CPU oracle (oracle/orc_cnn.py) pinned by what the
mathematics fixes: its analytic gradients against central finite
differences of its own float64 loss (every layer, weights and biases, the
head), the bf16 mode within bf16 rounding of the float64 one, the conv as a
direct sum, and the loss's nearest-hypothesis rule on hand-made cases."""
import numpy as np
import pytest

import orc_cnn as oc


def _case(seed, n=1):
    rs = np.random.default_rng(seed)
    p = oc.init_params(seed)
    p[:] += rs.normal(0, 0.01, p.shape).astype(np.float32)  # nonzero biases too
    x = oc.bf16(rs.normal(0, 0.6, (n, 64, 64, 2)).astype(np.float32))
    mvin = rs.normal(0, 0.5, (n, 16, 32))
    target = rs.integers(-20, 21, (n, 16, 2)).astype(np.float64)
    return p.astype(np.float64), x, mvin, target


def test_gradients_match_finite_differences():
    p, x, mvin, target = _case(3)
    acts, pool, z = oc.forward(p, x, mvin, rounding=None)
    _, dz = oc.loss_and_dz(z, target)
    g, _ = oc.backward(p, acts, pool, mvin, dz, rounding=None)
    rs = np.random.default_rng(4)
    idx = []
    for l, (cin, cout) in enumerate(oc.LAYERS):  # weights and the bias of every layer
        kp = oc.kpad(cin)
        for _ in range(6):
            idx.append(oc.OFFSETS[l] + int(rs.integers(cout)) * (kp + 1) + int(rs.integers(9 * cin)))
        idx.append(oc.OFFSETS[l] + int(rs.integers(cout)) * (kp + 1) + kp)
    idx += [oc.OFFSETS[4] + int(rs.integers(oc.NOUT * (oc.NFEAT + 1))) for _ in range(8)]
    h = 1e-6  # small enough that no ReLU or Huber kink lies within +-h
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = oc.total_loss(q, x, mvin, target)
        q[i] -= 2 * h
        lm = oc.total_loss(q, x, mvin, target)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 2e-5 * max(1.0, abs(num)), (i, num, g[i])


def test_padding_columns_get_no_gradient():
    p, x, mvin, target = _case(5)
    acts, pool, z = oc.forward(p, x, mvin, rounding=None)
    g, _ = oc.backward(p, acts, pool, mvin, oc.loss_and_dz(z, target)[1], rounding=None)
    for l, (cin, cout) in enumerate(oc.LAYERS):
        kp = oc.kpad(cin)
        m = g[oc.OFFSETS[l]:oc.OFFSETS[l] + cout * (kp + 1)].reshape(cout, kp + 1)
        assert not m[:, 9 * cin:kp].any()


def test_conv_is_the_direct_sum():
    rs = np.random.default_rng(6)
    x = rs.normal(size=(1, 64, 64, 3))
    w = rs.normal(size=(4, 3, 3, 3))
    b = rs.normal(size=4)
    y = oc.conv(x, w, b)
    for (yy, xx, o) in ((0, 0, 0), (10, 63, 3), (63, 31, 1), (5, 7, 2)):
        s = b[o]
        for ky in range(3):
            for kx in range(3):
                sy, sx = yy + ky - 1, xx + kx - 1
                if 0 <= sy < 64 and 0 <= sx < 64:
                    s += (w[o, ky, kx] * x[0, sy, sx]).sum()
        assert abs(s - y[0, yy, xx, o]) < 1e-9


def test_strided_conv_is_the_direct_sum():
    """Stride 2: output (y, x) sums the input window centred on (2y, 2x)."""
    rs = np.random.default_rng(8)
    x = rs.normal(size=(1, 32, 32, 3))
    w = rs.normal(size=(4, 3, 3, 3))
    b = rs.normal(size=4)
    y = oc.conv(x, w, b, stride=2)
    assert y.shape == (1, 16, 16, 4)
    for (yy, xx, o) in ((0, 0, 0), (15, 15, 3), (7, 0, 1), (3, 12, 2)):
        s = b[o]
        for ky in range(3):
            for kx in range(3):
                sy, sx = 2 * yy + ky - 1, 2 * xx + kx - 1
                if 0 <= sy < 32 and 0 <= sx < 32:
                    s += (w[o, ky, kx] * x[0, sy, sx]).sum()
        assert abs(s - y[0, yy, xx, o]) < 1e-9


def test_network_shapes_and_cells():
    """conv1 and conv2 halve the side, the cell sums are the 4x4 cells of
    conv4 in raster order (cell j covers the tile's 16x16 block j)."""
    p, x, mvin, _ = _case(9)
    acts, pool, _ = oc.forward(p, x, mvin, rounding=None)
    assert [a.shape[1:] for a in acts] == [(64, 64, 2), (32, 32, 16), (16, 16, 32), (16, 16, 32), (16, 16, 32)]
    a4 = acts[4]
    for j in (0, 5, 15):
        cy, cx = divmod(j, 4)
        np.testing.assert_allclose(pool[0, j], a4[0, 4 * cy:4 * cy + 4, 4 * cx:4 * cx + 4].sum(axis=(0, 1)))


def test_bf16_mode_stays_near_float64():
    p, x, mvin, _ = _case(7)
    _, _, z64 = oc.forward(p, x, mvin, rounding=None)
    _, _, z16 = oc.forward(p, x, mvin, rounding="bf16")
    assert np.abs(z16 - z64).max() < 0.05 * max(1.0, np.abs(z64).max())


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5 + 2 ** -9], np.float32)
    np.testing.assert_array_equal(oc.bf16(v), np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5], np.float32))


def test_loss_picks_the_nearest_hypothesis():
    z = np.zeros((1, 1, 12))
    z[0, 0, :8] = [0.5, 0.5, 0.1, 0.0, 0.1, 0.0, -1, -1]   # hypotheses 1 and 2 tie: the first wins
    loss, dz = oc.loss_and_dz(z, np.array([[[1.6, 0.0]]]))  # target (0.1, 0) in units of 16 pel
    assert dz[0, 0, 2] == 0 and dz[0, 0, 3] == 0            # the winner sits on the target
    assert dz[0, 0, 8 + 1] < 0 and dz[0, 0, 8 + 0] > 0     # its confidence is pushed up, others down
    assert abs(dz[0, 0, 4] - oc.EPS * 0.0) < 1e-12          # the tied loser: EPS times a zero gradient
    assert abs(dz[0, 0, 0] - oc.EPS * 1.0) < 1e-12          # a far one: EPS times the Huber slope 1
    assert abs(loss[0, 0] - (oc.EPS * (0.4 - 1 / 32 + 0.5 - 1 / 32 + 1.1 - 1 / 32 + 1.0 - 1 / 32) + np.log(4))) < 1e-9
