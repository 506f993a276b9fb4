// hevc_intra.hip - HEVC I-picture coding on gfx950 as ONE persistent
// wavefront launch: encode_ctu's I branch with intra_trial
// (enc/src/hevc_enc_template.c:502-594,1409-1427) for every CTU.
//
// Coding shape (fixed, as in the reference): CTU16 -> 4 CU8 -> 4 PU4 in
// z-order.  Each luma PU takes the SAD-best of Planar / DC / H / V (strict
// '<' in that order) built from reconstructed neighbours with z-scan
// availability and the substitution scan of the reference's gather
// (hevc_intra_template.c:28-112), then DST-VII, quantisation, dequantisation,
// inverse and reconstruction, so the next PU sees it.  Chroma is one 4x4 DC
// block per CU and component (DCT-II, QpC).
//
// Work layout: one wave per CTU row, 4 rows per workgroup (row group from an
// atomic ticket, so a row's producer always started before it).  Per CTU the
// wave runs 16 steps, one per luma PU: lanes 0..15 hold the PU's samples,
// lanes 16..31 / 32..47 the Cb / Cr block of the CU on its first step.  A
// CTU needs the bottom rows of CTUs x-1, x, x+1 of the row above (corner,
// top, top-right); each finished CTU hands its bottom luma / Cb / Cr row (32
// samples) to the wave below through an LDS ring, or to the next workgroup as
// system-coherent 64-bit records {tag = (epoch << 16) | (x + 1), one dword
// of samples} that the reader polls until every tag matches (cleared by
// system-scope stores before the launch; the launch epoch keeps the previous
// picture's record at the same place from ever matching).  Rows
// that start a slice read nothing above and wait for nothing.  Waits are
// bounded and report on *err.
//
// 8-bit NV12 or 10-bit P010 frames (the sample in the high 10 bits of each
// 16-bit word); inside the kernel samples are plain values of type Pix.
#include "bc250_hevc.h"
#include "hevc_dev.h"

using namespace hevc;

namespace {
constexpr int WAVES = 4;
constexpr int RING = 16;
constexpr int REC_MAX = 16;  // dwords per CTU record at 16 bits per sample
// Tiles: row 0 = the row above (col 3 corner, then the CTU's width, then 4
// top-right samples), rows 1.. = the CTU with col 3 the left neighbour column.
constexpr int LP = 24;   // luma tile pitch (cols 3..23), 17 rows
constexpr int CPT = 16;  // chroma tile pitch (cols 3..15), 9 rows

// LDS samples of one wave, addressed by offset so a lane can pick its plane.
constexpr int O_TY = 0;                    // luma tile 17 x LP
constexpr int O_TC = O_TY + 17 * LP;       // Cb, Cr tiles 9 x CPT each
constexpr int O_AY = O_TC + 2 * 9 * CPT;   // records of the row above, luma: 4 CTU slots x 16
constexpr int O_AC = O_AY + 4 * 16;        // records of the row above, Cb / Cr: 2 x 4 slots x 8
constexpr int O_SRC = O_AC + 2 * 4 * 8;    // source CTU: luma 256, Cb 64, Cr 64
constexpr int WAVE_SAMPLES = O_SRC + 384;

template <typename Pix>
struct alignas(16) WaveLds {
    Pix b[WAVE_SAMPLES];
};

// z-scan order of the 4x4 blocks inside a CTU (hevc_intra.c:75-130): the
// Morton code of the block coordinates (CTU16: 2 bits each, CTU8: 1 bit).
BC_DEV int morton4(int rx, int ry)
{
    const int bx = rx >> 2, by = ry >> 2;
    return ((by >> 1) << 3) | ((bx >> 1) << 2) | ((by & 1) << 1) | (bx & 1);
}

// Cross-lane moves inside each row of 16 lanes (one 4x4 block, lane =
// 4 * row + column): down_rows<t> gives every lane the value of the lane t
// block rows above it (row_ror by 4t, wrapping), next_cols<t> the value of the
// lane t columns to its right in the same block row (quad_perm rotation).
template <int T>
BC_DEV int down_rows(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x120 + 4 * T, 0xF, 0xF, false);
}

template <int T>
BC_DEV int next_cols(int v)
{
    constexpr int ctrl = T == 1 ? 0x39 : (T == 2 ? 0x4E : 0x93);  // [1,2,3,0], [2,3,0,1], [3,0,1,2]
    return __builtin_amdgcn_update_dpp(0, v, ctrl, 0xF, 0xF, false);
}

// One pass of a 4x4 transform over the block held by a row of 16 lanes, the
// lane's coefficients already permuted to match the moved values: columns
// (sum over block rows) or rows (sum over block columns).
// Values are int16 and |c| <= 90, so 24-bit multiplies (full rate) are exact.
BC_DEV int col_pass(int v, const int c[4])
{
    return __mul24(c[0], v) + __mul24(c[1], down_rows<1>(v)) + __mul24(c[2], down_rows<2>(v)) +
           __mul24(c[3], down_rows<3>(v));
}

BC_DEV int row_pass(int v, const int c[4])
{
    return __mul24(c[0], v) + __mul24(c[1], next_cols<1>(v)) + __mul24(c[2], next_cols<2>(v)) +
           __mul24(c[3], next_cols<3>(v));
}

// Sum over each row of 16 lanes, result in every lane of the row: DPP
// quad_perm [1,0,3,2] and [2,3,0,1], then half-row mirror and row mirror (all
// gfx9 DPP forms), so no LDS round trip.
BC_DEV uint32_t row16_sum(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
    return v;
}

}  // namespace

__global__ void hevc_clear_records_kernel(uint64_t *rec, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st_sys64(rec + i, 0);
}

template <typename Pix, int BD>
__global__ __launch_bounds__(256) void hevc_i_encode_kernel(
    const uint8_t *__restrict__ cy, const uint8_t *__restrict__ cuv, int cpy, int cpuv, uint8_t *__restrict__ oy,
    uint8_t *__restrict__ ouv, int opy, int opuv, int wc, int hc, int qp, int num_slices, int round_q12,
    uint8_t *__restrict__ modes, int16_t *__restrict__ coef, int *sync, uint64_t *ghand, unsigned spin_limit,
    int ctu_log2)
{
    // samples per dword; record dwords: luma row (LD), Cb row, Cr row (CD each)
    constexpr int SPD = 4 / (int)sizeof(Pix), LD = 16 / SPD, CD = 8 / SPD, REC = LD + 2 * CD;
    constexpr int MAXV = (1 << BD) - 1, QOFF = 6 * (BD - 8);
    constexpr int S1 = BD == 8 ? 1 : 3, QSH = BD - 3, S2INV = 20 - BD;
    __shared__ WaveLds<Pix> wl[WAVES];
    __shared__ uint32_t ring[WAVES - 1][RING][REC];
    __shared__ int prog[WAVES], cons[WAVES], done[WAVES];
    __shared__ int bls[WAVES][3];
    __shared__ int s_group;

    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    int *err = sync + 1;
    const uint32_t ep = wf_epoch((const uint32_t *)sync + 2);
    if (tid == 0) s_group = atomicAdd(sync, 1);
    if (tid < WAVES) prog[tid] = cons[tid] = done[tid] = 0;
    __syncthreads();
    const int wtu = wc >> 4, htu = hc >> 4;
    const int yrow = s_group * WAVES + wv;
    if (yrow >= htu) return;
    volatile int *vprog = prog, *vcons = cons, *vdone = done;
    Pix *B = wl[wv].b;

    const int r0 = slice_first_row16(yrow, htu, num_slices, ctu_log2);
    const bool need_above = yrow > r0;
    const bool give_below = yrow + 1 < htu && slice_first_row16(yrow + 1, htu, num_slices, ctu_log2) != yrow + 1;
    // CTU32: rows pair up (2r, 2r+1) in one workgroup and the four 16x16
    // quadrants of a CTU32 go top-left, top-right, bottom-left, bottom-right,
    // so the upper row also waits for the lower one: the top-left quadrant of
    // CTU32 j+1 has its bottom-left neighbour in the bottom-right one of j.
    const bool c32 = ctu_log2 == 5, lower = c32 && (yrow & 1);
    const bool has_lower = c32 && !(yrow & 1) && yrow + 1 < htu;
    const bool lds_above = wv > 0, lds_below = wv < WAVES - 1;

    // lane roles: group 0 luma PU samples, 1 Cb, 2 Cr (3: idle copy of Cr)
    const int grp = lane >> 4, k16 = lane & 15, pi = k16 >> 2, pj = k16 & 3;
    const bool luma = grp == 0;
    const int plane = grp > 2 ? 2 : grp;
    const int cs = luma ? 16 : 8;
    const int tile = luma ? O_TY : O_TC + (plane - 1) * 9 * CPT, tp = luma ? LP : CPT;
    const int srco = luma ? O_SRC : O_SRC + 256 + (plane - 1) * 64;
    // the matrix entries this lane multiplies by, in the order the moves of
    // col_pass / row_pass deliver the block: forward columns M[pi][pi-t],
    // forward rows M[pj][pj+t], inverse columns M[pi-t][pi], inverse rows
    // M[pj+t][pj] (indices mod 4)
    int fc[4], fr[4], ic[4], ir[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int8_t *M = luma ? kDst4 : kDct4;
        const int a = (pi - t) & 3, b = (pj + t) & 3;
        fc[t] = M[pi * 4 + a];
        fr[t] = M[pj * 4 + b];
        ic[t] = M[a * 4 + pi];
        ir[t] = M[b * 4 + pj];
    }
    const int qy = clampi(qp + QOFF, 0, BD == 8 ? 51 : 63), cqp = hevc::chroma_qp(qp) + QOFF;
    const Quant QL = make_quant(qy, QSH, round_q12), QC = make_quant(clampi(cqp, 0, 63), QSH, round_q12);
    // this lane's quantiser, computed here once (the empty asm keeps the
    // compiler from recomputing the division and 64-bit products per PU)
    Quant QP = luma ? QL : QC;
    asm("" : "+v"(QP.p), "+v"(QP.shift), "+v"(QP.d), "+v"(QP.off), "+v"(QP.m));
    const int q = luma ? qy : cqp;

    // source prefetch: 4 luma samples per lane, 4 interleaved chroma samples per lane 0..31
    const int Y0r = yrow * 16;
    int pre_y[4], pre_c[4] = {0, 0, 0, 0};
    auto prefetch = [&](int x) {
        ld4<Pix>(cy + (size_t)(Y0r + (lane >> 2)) * cpy, x * 16 + (lane & 3) * 4, pre_y);
        if (lane < 32) ld4<Pix>(cuv + (size_t)(yrow * 8 + (lane >> 2)) * cpuv, x * 16 + (lane & 3) * 4, pre_c);
    };
    prefetch(0);

    // CTU-local sample (rx in -1..cs+3, ry in -1..cs-1) of this lane's plane
    auto sample = [&](int rx, int ry) -> int { return B[tile + (ry + 1) * tp + 4 + rx]; };
    auto take_record = [&](int j, uint32_t v, int d) {
        // dword d of record j: luma row samples, then Cb, then Cr, SPD per dword
        if (d < LD) *(uint32_t *)&B[O_AY + (j & 3) * 16 + d * SPD] = v;
        else if (d < LD + CD) *(uint32_t *)&B[O_AC + (j & 3) * 8 + (d - LD) * SPD] = v;
        else *(uint32_t *)&B[O_AC + 32 + (j & 3) * 8 + (d - LD - CD) * SPD] = v;
    };
    auto fetch_above = [&](int j) {  // record j of the row above into its slot
        uint32_t v = 0;
        if (lds_above) {
            if (lane < REC) v = ring[wv - 1][j % RING][lane];
        } else {
            const uint64_t *src = ghand + ((size_t)(yrow - 1) * wtu + j) * REC + lane;
            unsigned n = 0;
            uint64_t qv = 0;
            for (;;) {
                bool ok = true;
                if (lane < REC) {
                    qv = ld_sys64(src);
                    ok = (uint32_t)(qv >> 32) == wf_tag(ep, j + 1);
                }
                if (__all(ok)) break;
                __builtin_amdgcn_s_sleep(1);
                if (++n > spin_limit) {
                    if (lane == 0) atomicOr(err, 1);
                    break;
                }
            }
            v = (uint32_t)qv;
        }
        if (lane < REC) take_record(j, v, lane);
    };

    int fetched = 0;  // records of the row above taken so far
    for (int x = 0; x < wtu; x++) {
        const int X0 = x * 16;
        // bottom-right quadrant: its top-right neighbour (the next CTU32) comes later
        const bool br = lower && (x & 1);
        // top-left quadrant after a CTU32: bottom-left neighbours in the lower row
        const bool tl_next = has_lower && !(x & 1) && x > 0;
        int blv = 0;
        if (tl_next) {
            unsigned n = 0;
            while (vdone[wv + 1] < x) {
                __builtin_amdgcn_s_sleep(1);
                if (++n > spin_limit) { atomicOr(err, 1); break; }
            }
            wave_lds_sync();
            blv = bls[wv + 1][plane];
        }
        // ---- records of the row above up to CTU x+1
        if (need_above) {
            const int need = br ? x + 1 : (x + 2 < wtu ? x + 2 : wtu);
            if (lds_above) {
                unsigned n = 0;
                while (vprog[wv - 1] < need) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++n > spin_limit) { atomicOr(err, 1); break; }
                }
                wave_lds_sync();
            }
            // records 0 .. need-1 (a CTU32 bottom-right quadrant takes none: the
            // next bottom-left one then takes two)
            while (fetched < need) fetch_above(fetched++);
            wave_lds_sync();
            if (lds_above && lane == 0) vcons[wv] = need;
            // tile row 0: corner (record x-1), top (record x), top-right (record x+1)
            if (lane < 21) {
                const int j = lane == 0 ? x - 1 : (lane <= 16 ? x : x + 1);
                const int o = lane == 0 ? 15 : (lane <= 16 ? lane - 1 : lane - 17);
                B[O_TY + 3 + lane] = B[O_AY + (j & 3) * 16 + o];
            } else if (lane < 47) {
                const int c = (lane - 21) / 13, t = (lane - 21) % 13;
                const int j = t == 0 ? x - 1 : (t <= 8 ? x : x + 1);
                const int o = t == 0 ? 7 : (t <= 8 ? t - 1 : t - 9);
                B[O_TC + c * 9 * CPT + 3 + t] = B[O_AC + c * 32 + (j & 3) * 8 + o];
            }
            wave_lds_sync();
        }
        // ---- source of CTU x into LDS; prefetch CTU x+1
#pragma unroll
        for (int i = 0; i < 4; i++) B[O_SRC + lane * 4 + i] = (Pix)pre_y[i];
        if (lane < 32) {
            const int r = lane >> 2, c2 = (lane & 3) * 2;
            B[O_SRC + 256 + r * 8 + c2] = (Pix)pre_c[0];
            B[O_SRC + 320 + r * 8 + c2] = (Pix)pre_c[1];
            B[O_SRC + 256 + r * 8 + c2 + 1] = (Pix)pre_c[2];
            B[O_SRC + 320 + r * 8 + c2 + 1] = (Pix)pre_c[3];
        }
        if (x + 1 < wtu) prefetch(x + 1);
        wave_lds_sync();

        const int ctx0 = luma ? X0 : X0 >> 1, cty0 = luma ? Y0r : Y0r >> 1;
        for (int k = 0; k < 16; k++) {
            const int cu = k >> 2, pu = k & 3;
            const bool active = luma || (grp < 3 && pu == 0);
            // ---- availability (z-scan inside the CTU, slices and edges outside),
            // written with bitwise logic and selects so that it stays branch-free
            const int rx = luma ? (cu & 1) * 8 + (pu & 1) * 4 : (cu & 1) * 4;
            const int ry = luma ? (cu >> 1) * 8 + (pu >> 1) * 4 : (cu >> 1) * 4;
            const int bx = ctx0 + rx, by = cty0 + ry, mz = morton4(rx, ry);
            const bool leftC = x > 0, aboveC = need_above, aboveR = need_above & (x + 1 < wtu) & !br;
            const bool in_r = rx + 4 < cs, in_b = ry + 4 < cs;
            // (the Morton codes of out-of-CTU positions are computed but never selected)
            const bool mtr = morton4(rx + 4, ry - 4) < mz, mbl = morton4(rx - 4, ry + 4) < mz;
            const bool aL = (rx > 0) | leftC;
            const bool aT = (ry > 0) | aboveC;
            const bool aC = rx > 0 ? ((ry > 0) | aboveC) : (ry > 0 ? leftC : (leftC & aboveC));
            const bool aTR = ry == 0 ? (in_r ? aboveC : aboveR) : (in_r & mtr);
            const bool blx = tl_next & (rx == 0) & (ry + 4 == cs);  // from the lower row's last quadrant
            const bool aBL = (in_b & (rx == 0 ? leftC : mbl)) | blx;
            // ---- neighbours with the substitution scan
            int v[10];
            bool a[10];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                a[t] = aL;
                v[t] = sample(rx - 1, ry + 3 - t);
                a[5 + t] = aT;
                v[5 + t] = sample(rx + t, ry - 1);
            }
            a[4] = aC;
            v[4] = sample(rx - 1, ry - 1);
            a[9] = aTR;
            v[9] = sample(rx + 4, ry - 1);
            const int vbs = sample(rx - 1, in_b ? ry + 4 : ry + 3);
            const int vbl = blx ? blv : vbs;
            // first available in scan order fills what precedes it; later gaps copy the previous
            int firstv = 1 << (BD - 1);
#pragma unroll
            for (int t = 9; t >= 0; t--)
                if (a[t]) firstv = v[t];
            int prev = firstv;
#pragma unroll
            for (int t = 0; t < 10; t++) {
                v[t] = a[t] ? v[t] : prev;
                prev = v[t];
            }
            const int L0 = v[3], L1 = v[2], L2 = v[1], L3 = v[0], Cc = v[4];
            const int T0 = v[5], T1 = v[6], T2 = v[7], T3 = v[8], T4 = v[9];
            const int L4 = aBL ? vbl : L3;
            const int Li = pi == 0 ? L0 : pi == 1 ? L1 : pi == 2 ? L2 : L3;
            const int Tj = pj == 0 ? T0 : pj == 1 ? T1 : pj == 2 ? T2 : T3;
            // ---- the four predictions of this lane's sample
            const int p_pl =
                (__mul24(3 - pj, Li) + __mul24(pj + 1, T4) + __mul24(3 - pi, Tj) + __mul24(pi + 1, L4) + 4) >> 3;
            const int dc = (L0 + L1 + L2 + L3 + T0 + T1 + T2 + T3 + 4) >> 3;
            // luma DC edge filter: corner, top row, left column
            const int dc_c = (L0 + 2 * dc + T0 + 2) >> 2, dc_t = (Tj + 3 * dc + 2) >> 2;
            const int dc_l = (Li + 3 * dc + 2) >> 2;
            const int dcf = pi == 0 ? (pj == 0 ? dc_c : dc_t) : (pj == 0 ? dc_l : dc);
            const int p_dc = luma ? dcf : dc;
            const int p_h = (luma && pi == 0) ? clampi(L0 + ((Tj - Cc) >> 1), 0, MAXV) : Li;
            const int p_v = (luma && pj == 0) ? clampi(T0 + ((Li - Cc) >> 1), 0, MAXV) : Tj;
            const int s = B[srco + (by - cty0 + pi) * cs + (bx - ctx0 + pj)];
            // ---- SAD decision over the 16 lanes of the group (luma only)
            // two SADs per word: 16 samples of at most 1023 stay below 2^16
            const uint32_t spd = row16_sum((uint32_t)iabs(s - p_pl) | (uint32_t)iabs(s - p_dc) << 16);
            const uint32_t shv = row16_sum((uint32_t)iabs(s - p_h) | (uint32_t)iabs(s - p_v) << 16);
            const uint32_t sp = spd & 0xffffu, sd = spd >> 16, sh = shv & 0xffffu, sv = shv >> 16;
            int mode = 0, pred = p_pl;
            uint32_t best = sp;
            if (sd < best) { best = sd; mode = 1; pred = p_dc; }
            if (sh < best) { best = sh; mode = 10; pred = p_h; }
            if (sv < best) { best = sv; mode = 26; pred = p_v; }
            if (!luma) { mode = 1; pred = dc; }
            // ---- forward transform, quantisation (cross-lane moves, no LDS)
            const int t1 = sat16((col_pass(s - pred, fc) + (1 << (S1 - 1))) >> S1);
            const int lv = quant(sat16((row_pass(t1, fr) + 128) >> 8), QP);
            // ---- dequantisation, inverse, reconstruction
            const int t2 = sat16((col_pass(dequant(lv, q, QSH), ic) + 64) >> 7);
            const int rec = clampi(pred + sat16((row_pass(t2, ir) + (1 << (S2INV - 1))) >> S2INV), 0, MAXV);
            if (active) {
                B[tile + (ry + pi + 1) * tp + 4 + (rx + pj)] = (Pix)rec;
                const size_t cb = (size_t)(yrow * wtu + x) * 384;
                if (luma) coef[cb + k * 16 + k16] = (int16_t)lv;
                else coef[cb + 256 + (plane - 1) * 64 + cu * 16 + k16] = (int16_t)lv;
                if (lane == 0) modes[(size_t)(yrow * wtu + x) * 16 + k] = (uint8_t)mode;
            }
            wave_lds_sync();
        }
        if (lower) {
            // bottom-right quadrant: its top-right samples are the bottom-left
            // neighbours of the next CTU32's top-left quadrant
            if ((x & 1) && lane < 3)
                bls[wv][lane] = lane == 0 ? B[O_TY + LP + 19] : B[O_TC + (lane - 1) * 9 * CPT + CPT + 11];
            wave_lds_sync();
            if (lane == 0) vdone[wv] = x + 1;
        }
        // ---- hand the CTU's bottom rows down
        if (give_below) {
            uint32_t h = 0;
            if (lane < LD) h = *(const uint32_t *)&B[O_TY + 16 * LP + 4 + lane * SPD];
            else if (lane < LD + CD) h = *(const uint32_t *)&B[O_TC + 8 * CPT + 4 + (lane - LD) * SPD];
            else if (lane < REC) h = *(const uint32_t *)&B[O_TC + 9 * CPT + 8 * CPT + 4 + (lane - LD - CD) * SPD];
            if (lds_below) {
                unsigned n = 0;
                while (vcons[wv + 1] < x + 1 - RING + 1) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++n > spin_limit) { atomicOr(err, 2); break; }
                }
                if (lane < REC) ring[wv][x % RING][lane] = h;
                wave_lds_sync();
                if (lane == 0) vprog[wv] = x + 1;
            } else if (lane < REC) {
                st_sys64(ghand + ((size_t)yrow * wtu + x) * REC + lane, ((uint64_t)wf_tag(ep, x + 1) << 32) | h);
            }
        }
        // ---- the CTU's reconstruction to the frame
        {
            const Pix *t = &B[O_TY + ((lane >> 2) + 1) * LP + 4 + (lane & 3) * 4];
            const int v[4] = {t[0], t[1], t[2], t[3]};
            st4<Pix>(oy + (size_t)(Y0r + (lane >> 2)) * opy, X0 + (lane & 3) * 4, v);
        }
        if (lane < 32) {
            const int r = lane >> 2, c2 = (lane & 3) * 2;
            const Pix *tb = &B[O_TC + (r + 1) * CPT + 4 + c2], *tr = &B[O_TC + 9 * CPT + (r + 1) * CPT + 4 + c2];
            const int v[4] = {tb[0], tr[0], tb[1], tr[1]};
            st4<Pix>(ouv + (size_t)(yrow * 8 + r) * opuv, X0 + (lane & 3) * 4, v);
        }
        wave_lds_sync();
        // ---- slide: this CTU's right column becomes the left neighbour column
        if (lane < 16) B[O_TY + (lane + 1) * LP + 3] = B[O_TY + (lane + 1) * LP + 19];
        else if (lane < 32) {
            const int c = (lane - 16) >> 3, r = (lane & 7) + 1;
            B[O_TC + c * 9 * CPT + r * CPT + 3] = B[O_TC + c * 9 * CPT + r * CPT + 11];
        }
        wave_lds_sync();
    }
}

extern "C" size_t bc250_hevc_intra_work_bytes(int wc, int hc)
{
    const size_t nctu = (size_t)(wc / 16) * (hc / 16);
    return 256 + nctu * REC_MAX * 8;
}

template <typename Pix, int BD>
static void launch_i(const uint8_t *cy, const uint8_t *cuv, int cpy, int cpuv, uint8_t *oy, uint8_t *ouv, int opy,
                     int opuv, int wc, int hc, int qp, int num_slices, int round_q12, uint8_t *modes, int16_t *coef,
                     int *sync, unsigned spin_limit, int ctu_log2, hipStream_t st)
{
    constexpr int REC = 32 / (4 / (int)sizeof(Pix));
    const int htu = hc / 16, nrec = (wc / 16) * htu * REC;
    uint64_t *ghand = (uint64_t *)((uint8_t *)sync + 256);  // after 256 B of flags: ticket, error, epoch
    wavefront_begin_kernel<<<1, 64, 0, st>>>(sync, 2, (uint32_t *)sync + 2);
    hevc_clear_records_kernel<<<(nrec + 255) / 256, 256, 0, st>>>(ghand, nrec);
    hevc_i_encode_kernel<Pix, BD><<<(htu + WAVES - 1) / WAVES, 64 * WAVES, 0, st>>>(
        cy, cuv, cpy, cpuv, oy, ouv, opy, opuv, wc, hc, qp, num_slices, round_q12, modes, coef, sync, ghand,
        spin_limit, ctu_log2);
}

extern "C" int bc250_hevc_launch_i_encode_ctu(const void *cy, const void *cuv, int cpy, int cpuv, void *oy,
                                              void *ouv, int opy, int opuv, int wc, int hc, int qp, int num_slices,
                                              int round_q12, uint8_t *modes, int16_t *coef, int *sync,
                                              unsigned spin_limit, int ctu_log2, int bd, void *stream)
{
    const int b = bd == 8 ? 1 : 2, al = 4 * b - 1;
    if ((bd != 8 && bd != 10) || (ctu_log2 != 4 && ctu_log2 != 5) || wc <= 0 || hc <= 0 || (wc & 15) ||
        (hc & 15) || qp < 0 || qp > 51 || wc / 16 >= 65535 ||
        num_slices < 1 || round_q12 < 0 || round_q12 > 12 || (cpy & al) || (cpuv & al) || (opy & al) ||
        (opuv & al) || cpy < wc * b || cpuv < wc * b || opy < wc * b || opuv < wc * b)
        return -1;
    hipStream_t st = (hipStream_t)stream;
    auto *C = (const uint8_t *)cy, *CU = (const uint8_t *)cuv;
    auto *O = (uint8_t *)oy, *OU = (uint8_t *)ouv;
    if (bd == 8)
        launch_i<uint8_t, 8>(C, CU, cpy, cpuv, O, OU, opy, opuv, wc, hc, qp, num_slices, round_q12, modes, coef, sync,
                             spin_limit, ctu_log2, st);
    else
        launch_i<uint16_t, 10>(C, CU, cpy, cpuv, O, OU, opy, opuv, wc, hc, qp, num_slices, round_q12, modes, coef,
                               sync, spin_limit, ctu_log2, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int bc250_hevc_launch_i_encode(const void *cy, const void *cuv, int cpy, int cpuv, void *oy, void *ouv,
                                          int opy, int opuv, int wc, int hc, int qp, int num_slices, int round_q12,
                                          uint8_t *modes, int16_t *coef, int *sync, unsigned spin_limit, int bd,
                                          void *stream)
{
    return bc250_hevc_launch_i_encode_ctu(cy, cuv, cpy, cpuv, oy, ouv, opy, opuv, wc, hc, qp, num_slices, round_q12,
                                          modes, coef, sync, spin_limit, 4, bd, stream);
}
