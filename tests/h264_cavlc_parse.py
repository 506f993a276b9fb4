"""A CAVLC slice parser for the syntax the H.264 slice writers emit (test
helper, not a test module).

Written from H.264 7.3.3 (slice header), 7.3.4 (slice_data), 7.3.5
(macroblock_layer, mb_pred, residual), 9.1 (Exp-Golomb), 9.2 (residual
block CAVLC: coeff_token, level_prefix / level_suffix, total_zeros,
run_before), 9.2.1 (nC from the neighbouring blocks' TotalCoeff), 8.4.1.1
(P_Skip motion) and 8.4.1.3 (median prediction, including the rule that a
lone neighbour with the same reference index is the prediction).  Only the
variable-length code tables are taken from the oracle's source text: they are
pinned entry by entry against the reference's own writer elsewhere, so what
this parser checks is everything around them.
"""
import os
import re

import numpy as np

ZZ = [0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15]


def _tables():
    src = open(os.path.join(os.path.dirname(__file__), "..", "oracle", "orc_cavlc.c")).read()

    def block(name):
        m = re.search(r"static const char \*const " + name + r"\[[^=]*= \{(.*?)\n\};", src, re.S)
        return [re.findall(r'"([01]*)"', row) for row in re.findall(r"\{([^{}]*)\}", m.group(1))]

    ct = block("CT")
    cts = []
    for t in range(3):
        d = {}
        for tc in range(17):
            for t1, code in enumerate(ct[t * 17 + tc]):
                if code:
                    d[code] = (tc, t1)
        cts.append(d)
    cdc = {}
    for tc, row in enumerate(block("CT_CDC")):
        for t1, code in enumerate(row):
            if code:
                cdc[code] = (tc, t1)
    tz = [{c: z for z, c in enumerate(row)} for row in block("TZ")]
    tzc = [{c: z for z, c in enumerate(row)} for row in block("TZ_CDC")]
    rb = [{c: r for r, c in enumerate(row)} for row in block("RB")]
    m = re.search(r"static const uint8_t tab\[48\] = \{(.*?)\};", src, re.S)
    inter = [int(x) for x in re.findall(r"\d+", m.group(1))]
    return cts, cdc, tz, tzc, rb, inter


_CT, _CDC, _TZ, _TZC, _RB, _CBP_INTER = _tables()


class Bits:
    def __init__(self, data):
        self.b = np.unpackbits(np.frombuffer(bytes(data), np.uint8))
        self.i = 0

    def u(self, n):
        v = 0
        for _ in range(n):
            if self.i >= len(self.b):
                raise ValueError("read past the end of the slice")
            v = v << 1 | int(self.b[self.i])
            self.i += 1
        return v

    def ue(self):
        z = 0
        while self.u(1) == 0:
            z += 1
            if z > 32:
                raise ValueError("Exp-Golomb prefix too long")
        return (1 << z) - 1 + self.u(z)

    def se(self):
        k = self.ue()
        return (k + 1) // 2 if k & 1 else -(k // 2)

    def vlc(self, table):
        code = ""
        while code not in table:
            code += str(self.u(1))
            if len(code) > 16:
                raise ValueError("no such codeword")
        return table[code]


def residual_block(bs, nc, maxc):
    """9.2: returns (coefficients in scan order, TotalCoeff)."""
    if nc == -1:
        tc, t1 = bs.vlc(_CDC)
    elif nc >= 8:
        v = bs.u(6)
        tc, t1 = (0, 0) if v == 3 else ((v >> 2) + 1, v & 3)
    else:
        tc, t1 = bs.vlc(_CT[0 if nc < 2 else 1 if nc < 4 else 2])
    coef = [0] * maxc
    if tc == 0:
        return coef, 0
    if tc > maxc or t1 > tc:
        raise ValueError("bad coeff_token")
    level = []
    for _ in range(t1):
        level.append(-1 if bs.u(1) else 1)
    sl = 1 if tc > 10 and t1 < 3 else 0
    for i in range(tc - t1):
        prefix = 0
        while bs.u(1) == 0:
            prefix += 1
        if prefix == 14 and sl == 0:
            size = 4
        elif prefix >= 15:
            size = prefix - 3
        else:
            size = sl
        code = (min(15, prefix) << sl) + (bs.u(size) if size else 0)
        if prefix >= 15 and sl == 0:
            code += 15
        if prefix >= 16:
            code += (1 << (prefix - 3)) - 4096
        if i == 0 and t1 < 3:
            code += 2
        lv = (code + 2) >> 1 if code % 2 == 0 else (-code - 1) >> 1
        level.append(lv)
        if sl == 0:
            sl = 1
        if abs(lv) > (3 << (sl - 1)) and sl < 6:
            sl += 1
    if tc < maxc:
        tzero = bs.vlc(_TZC[tc - 1] if nc == -1 else _TZ[tc - 1])
    else:
        tzero = 0
    runs, left = [], tzero
    for i in range(tc - 1):
        r = bs.vlc(_RB[min(left, 7) - 1]) if left > 0 else 0
        runs.append(r)
        left -= r
    runs.append(left)
    pos = -1
    for i in range(tc - 1, -1, -1):
        pos += runs[i] + 1
        coef[pos] = level[i]
    return coef, tc


def parse_sps(rbsp):
    """seq_parameter_set_data (7.3.2.1.1) for the subset the writer emits:
    no scaling matrices, pic_order_cnt_type 0 or 2, VUI with an aspect
    ratio and optional timing."""
    bs = Bits(rbsp)
    d = {"profile_idc": bs.u(8), "constraint_flags": bs.u(8), "level_idc": bs.u(8), "sps_id": bs.ue()}
    if d["profile_idc"] in (100, 110, 122, 244, 44, 83, 86, 118, 128):
        d["chroma_format_idc"] = bs.ue()
        if d["chroma_format_idc"] == 3:
            d["separate_colour_plane"] = bs.u(1)
        d["bit_depth_luma"] = bs.ue() + 8
        d["bit_depth_chroma"] = bs.ue() + 8
        d["transform_bypass"] = bs.u(1)
        if bs.u(1):
            raise ValueError("seq_scaling_matrix_present: not emitted by the writer")
    else:
        d["chroma_format_idc"], d["bit_depth_luma"], d["bit_depth_chroma"] = 1, 8, 8
    d["log2_max_frame_num_minus4"] = bs.ue()
    d["pic_order_cnt_type"] = bs.ue()
    if d["pic_order_cnt_type"] == 0:
        d["log2_max_poc_lsb_minus4"] = bs.ue()
    elif d["pic_order_cnt_type"] == 1:
        raise ValueError("pic_order_cnt_type 1: not emitted by the writer")
    d["max_num_ref_frames"] = bs.ue()
    d["gaps_allowed"] = bs.u(1)
    d["width_in_mbs"] = bs.ue() + 1
    d["height_in_map_units"] = bs.ue() + 1
    d["frame_mbs_only"] = bs.u(1)
    if not d["frame_mbs_only"]:
        d["mb_adaptive_frame_field"] = bs.u(1)
    d["direct_8x8_inference"] = bs.u(1)
    d["crop"] = [0, 0, 0, 0]
    if bs.u(1):
        d["crop"] = [bs.ue() for _ in range(4)]
    d["vui_present"] = bs.u(1)
    if d["vui_present"]:
        if bs.u(1):
            idc = bs.u(8)
            d["sar"] = (bs.u(16), bs.u(16)) if idc == 255 else idc
        if bs.u(1) or bs.u(1) or bs.u(1):
            raise ValueError("overscan / video signal / chroma location info: not emitted by the writer")
        if bs.u(1):
            d["num_units_in_tick"], d["time_scale"], d["fixed_frame_rate"] = bs.u(32), bs.u(32), bs.u(1)
        if bs.u(1) or bs.u(1) or bs.u(1) or bs.u(1):
            raise ValueError("HRD / pic_struct / bitstream restriction: not emitted by the writer")
    if bs.u(1) != 1:
        raise ValueError("SPS: no rbsp stop bit")
    return d


def parse_pps(rbsp):
    """pic_parameter_set_rbsp (7.3.2.2), one slice group."""
    bs = Bits(rbsp)
    d = {"pps_id": bs.ue(), "sps_id": bs.ue(), "entropy_coding_mode": bs.u(1), "pic_order_present": bs.u(1)}
    if bs.ue() != 0:
        raise ValueError("slice groups: not emitted by the writer")
    d["num_ref_idx_l0_default_minus1"] = bs.ue()
    d["num_ref_idx_l1_default_minus1"] = bs.ue()
    d["weighted_pred"] = bs.u(1)
    d["weighted_bipred_idc"] = bs.u(2)
    d["pic_init_qp_minus26"] = bs.se()
    d["pic_init_qs_minus26"] = bs.se()
    d["chroma_qp_offset"] = bs.se()
    d["deblocking_filter_control"] = bs.u(1)
    d["constrained_intra_pred"] = bs.u(1)
    d["redundant_pic_cnt_present"] = bs.u(1)
    return d


def _median(a, b, c):
    return sorted((a, b, c))[1]


class SliceParser:
    """Parses every slice of one frame.  Results per macroblock address:
    kind ('I', 'P', 'skip'), pm / cm (I16x16 modes), cbp, mv, the 24 x 16
    levels in the writer's raster layout and the I16x16 luma DC / chroma DC
    levels."""

    def __init__(self, wmb, hmb, num_slices, log2_max_frame_num_minus4=0, log2_max_poc_lsb_minus4=0):
        self.wmb, self.hmb, self.ns = wmb, hmb, num_slices
        self.fn_bits = log2_max_frame_num_minus4 + 4
        self.poc_bits = log2_max_poc_lsb_minus4 + 4
        n = wmb * hmb
        self.kind = [None] * n
        self.pm = np.zeros(n, int)
        self.cm = np.zeros(n, int)
        self.cbp = np.zeros(n, int)
        self.mv = np.zeros((n, 2), int)
        self.levels = np.zeros((n, 24, 16), int)
        self.ldc = np.zeros((n, 16), int)
        self.cdc = np.zeros((n, 2, 4), int)
        self.tcl = np.zeros((n, 16), int)     # TotalCoeff per luma blkIdx
        self.tcc = np.zeros((n, 2, 4), int)
        self.slice_of = np.full(n, -1)
        self.headers = []

    # -- neighbours (6.4.9 with slices) --
    def _avail(self, mb, cur):
        return 0 <= mb < len(self.kind) and self.slice_of[mb] == self.slice_of[cur] and self.kind[mb] is not None

    def _nb(self, mb, dx, dy):
        x, y = mb % self.wmb + dx, mb // self.wmb + dy
        if x < 0 or y < 0 or x >= self.wmb:
            return -1
        n = y * self.wmb + x
        return n if self._avail(n, mb) else -1

    @staticmethod
    def _blk(x, y):
        return (x & 1) | (y & 1) << 1 | (x & 2) << 1 | (y & 2) << 2

    def _nc_luma(self, mb, k):
        x, y = (k & 1) | (k & 4) >> 1, (k & 2) >> 1 | (k & 8) >> 2
        na = nb = None
        if x > 0:
            na = self.tcl[mb][self._blk(x - 1, y)]
        elif (a := self._nb(mb, -1, 0)) >= 0:
            na = self.tcl[a][self._blk(3, y)]
        if y > 0:
            nb = self.tcl[mb][self._blk(x, y - 1)]
        elif (b := self._nb(mb, 0, -1)) >= 0:
            nb = self.tcl[b][self._blk(x, 3)]
        return self._nc(na, nb)

    def _nc_chroma(self, mb, comp, k):
        x, y = k & 1, k >> 1
        na = nb = None
        if x > 0:
            na = self.tcc[mb][comp][k - 1]
        elif (a := self._nb(mb, -1, 0)) >= 0:
            na = self.tcc[a][comp][k + 1]
        if y > 0:
            nb = self.tcc[mb][comp][k - 2]
        elif (b := self._nb(mb, 0, -1)) >= 0:
            nb = self.tcc[b][comp][k + 2]
        return self._nc(na, nb)

    @staticmethod
    def _nc(na, nb):
        if na is not None and nb is not None:
            return (na + nb + 1) >> 1
        return na if na is not None else nb if nb is not None else 0

    # -- motion (8.4.1.3, 8.4.1.1), one reference picture --
    def _pred(self, mb):
        a, b = self._nb(mb, -1, 0), self._nb(mb, 0, -1)
        c = self._nb(mb, 1, -1)
        if c < 0:
            c = self._nb(mb, -1, -1)
        if b < 0 and c < 0 and a >= 0:
            b = c = a
        have = [n for n in (a, b, c) if n >= 0]
        if len(have) == 1:
            return tuple(self.mv[have[0]])
        v = [self.mv[n] if n >= 0 else (0, 0) for n in (a, b, c)]
        return _median(v[0][0], v[1][0], v[2][0]), _median(v[0][1], v[1][1], v[2][1])

    def _skip_mv(self, mb):
        a, b = self._nb(mb, -1, 0), self._nb(mb, 0, -1)
        if a < 0 or b < 0 or tuple(self.mv[a]) == (0, 0) or tuple(self.mv[b]) == (0, 0):
            return 0, 0
        return self._pred(mb)

    def _chroma(self, bs, mb, cbp_c):
        if cbp_c >= 1:
            for comp in range(2):
                c, _ = residual_block(bs, -1, 4)
                self.cdc[mb][comp] = c
        if cbp_c == 2:
            for comp in range(2):
                for k in range(4):
                    c, tc = residual_block(bs, self._nc_chroma(mb, comp, k), 15)
                    for i in range(15):
                        self.levels[mb][16 + comp * 4 + k][ZZ[i + 1]] = c[i]
                    self.tcc[mb][comp][k] = tc

    def _mb_i16(self, bs, mb, t):
        if not 1 <= t <= 24:
            raise ValueError(f"mb {mb}: mb_type {t} is not I_16x16")
        self.kind[mb] = "I"
        self.pm[mb], cbp_c, ac = (t - 1) % 4, (t - 1) // 4 % 3, (t - 1) // 12
        self.cm[mb] = bs.ue()
        if bs.se() != 0:
            raise ValueError("mb_qp_delta != 0")
        self.cbp[mb] = cbp_c << 4 | (15 if ac else 0)
        c, _ = residual_block(bs, self._nc_luma(mb, 0), 16)
        for i in range(16):
            self.ldc[mb][ZZ[i]] = c[i]
        if ac:
            for k in range(16):
                c, tc = residual_block(bs, self._nc_luma(mb, k), 15)
                r = ((k & 2) >> 1 | (k & 8) >> 2) * 4 + ((k & 1) | (k & 4) >> 1)
                for i in range(15):
                    self.levels[mb][r][ZZ[i + 1]] = c[i]
                self.tcl[mb][k] = tc
        self._chroma(bs, mb, cbp_c)

    def _mb_p(self, bs, mb):
        t = bs.ue()
        if t != 0:
            raise ValueError(f"mb {mb}: mb_type {t} is not P_L0_16x16")
        self.kind[mb] = "P"
        mvd = bs.se(), bs.se()
        px, py = self._pred(mb)
        self.mv[mb] = px + mvd[0], py + mvd[1]
        cbp = _CBP_INTER[bs.ue()]
        self.cbp[mb] = cbp
        if cbp:
            if bs.se() != 0:
                raise ValueError("mb_qp_delta != 0")
        for k in range(16):
            if cbp >> (k >> 2) & 1:
                c, tc = residual_block(bs, self._nc_luma(mb, k), 16)
                r = ((k & 2) >> 1 | (k & 8) >> 2) * 4 + ((k & 1) | (k & 4) >> 1)
                for i in range(16):
                    self.levels[mb][r][ZZ[i]] = c[i]
                self.tcl[mb][k] = tc
        self._chroma(bs, mb, cbp >> 4)

    def parse(self, s, rbsp):
        total = self.wmb * self.hmb
        start, end = s * total // self.ns, (s + 1) * total // self.ns
        bs = Bits(rbsp)
        h = {"first_mb": bs.ue(), "slice_type": bs.ue(), "pps_id": bs.ue(), "frame_num": bs.u(self.fn_bits)}
        idr = h["slice_type"] == 2
        if idr:
            h["idr_pic_id"] = bs.ue()
        h["poc_lsb"] = bs.u(self.poc_bits)
        if not idr:
            h["num_ref_idx_override"] = bs.u(1)
            h["ref_pic_list_mod"] = bs.u(1)
            h["adaptive_marking"] = bs.u(1)
        else:
            h["no_output_of_prior_pics"] = bs.u(1)
            h["long_term_reference"] = bs.u(1)
        h["slice_qp_delta"] = bs.se()
        h["deblock_idc"] = bs.ue()
        if h["deblock_idc"] != 1:
            h["alpha"], h["beta"] = bs.se(), bs.se()
        if h["first_mb"] != start:
            raise ValueError(f"slice {s}: first_mb {h['first_mb']} != {start}")
        self.headers.append(h)
        self.slice_of[start:end] = s
        mb = start
        while mb < end:
            if idr:
                self._mb_i16(bs, mb, bs.ue())
                mb += 1
                continue
            run = bs.ue()
            if mb + run > end:
                raise ValueError("mb_skip_run past the slice end")
            for _ in range(run):
                self.kind[mb] = "skip"
                self.mv[mb] = self._skip_mv(mb)
                mb += 1
            if mb < end:
                self._mb_p(bs, mb)
                mb += 1
        # rbsp_slice_trailing_bits: a stop bit, then zeros to the byte end
        if bs.u(1) != 1:
            raise ValueError(f"slice {s}: no stop bit after the last macroblock")
        rest = bs.b[bs.i:]
        if len(rest) >= 8 or rest.any():
            raise ValueError(f"slice {s}: {len(rest)} bits after the stop bit")
        return h
