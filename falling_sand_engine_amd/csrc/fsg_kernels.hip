// fsg_kernels.hip -- HIP kernels of the falling-sand tick for gfx950.
//
// k_tick_world / k_tick_batch: one workgroup updates one 64x64 chunk for one
// whole tick.  It stages the chunk plus a 4-cell halo (72x72 cells) in LDS as
// 32-bit compact cells (the stored low word, with the cell's origin and a
// reaction mark in its spare bits: fsg_device.hpp) and keeps the colours
// (the high words) in registers.  It runs every Margolus phase of the tick
// and the temperature step on the LDS copy (blocks are disjoint inside a
// phase, so no atomics are needed; halo cells are recomputed redundantly and
// shrink by one ring per phase), then lays the colours out in LDS, gathers
// each output cell's colour from its origin and writes the chunk interior
// once.  About 24 KB of LDS per workgroup keeps several workgroups resident
// per CU, so one workgroup's loads and stores overlap the others' phases.
// HBM traffic is one read and one write of each cell per tick plus the halo
// re-read that the L2/MALL mostly absorbs.  This replaces the reference's
// 16-thread chunk pool (world.cpp:1057-1077) and its 3 sequential passes per
// chunk.
#include "fsg_device.hpp"
#include "fsg_kernels.hpp"

namespace fsg {

static_assert(NT == 256, "one pair-table entry per thread and iteration");
static_assert(TW == 64 && TH == 64 && HALO == 4, "thread layout of the write-back");

__constant__ int kOX[4] = {0, 1, 1, 0};
__constant__ int kOY[4] = {0, 1, 0, 1};

__device__ __forceinline__ bool wood_hot(const MatTab &M, uint32_t c) {
    return cbyte(c) == M.wood && ctemp(c) > M.burn;
}

// Bands touched by a change covering chunk cells [x0, x1] x [y0, y1]
// (chunk coordinates, already clipped to the chunk).
__device__ __forceinline__ uint32_t band_bits(int x0, int x1, int y0, int y1) {
    const bool n = y0 < HALO, so = y1 >= TH - HALO, we = x0 < HALO, ea = x1 >= TW - HALO;
    return F_ANY | (n ? F_N : 0u) | (so ? F_S : 0u) | (we ? F_W : 0u) | (ea ? F_E : 0u) |
           ((n && we) ? F_NW : 0u) | ((n && ea) ? F_NE : 0u) | ((so && we) ? F_SW : 0u) |
           ((so && ea) ? F_SE : 0u);
}

// Temperature of a neighbour for the 5-point step: a WALL cell (outside the
// world) counts as the cell itself.  WALLS = false where the region has none.
template <bool WALLS>
__device__ __forceinline__ int ntemp(int T, uint32_t n) {
    return (WALLS && cid(n) == WALL) ? T : ctemp(n);
}

// New temperature of a cell with temperature T, material id m and the given
// neighbour temperatures: integer diffusion with coefficient 1/8 and floor
// rounding, plus the material's add_temp (addt, in LDS), clamped to int16.
__device__ __forceinline__ int temp_of(const MatTab &M, const int32_t *addt, int T, uint32_t m, int tu, int td,
                                       int tl, int tr) {
    int v = T + ((tu + td + tl + tr - 4 * T) >> 3);
    if (M.any_add_temp) {
        v += addt[m];
        v = min(max(v, -32768), 32767);
    }
    return v;
}

// Dirty-rect predicate over the staged tile, from the pair table of
// iteration 0 (where every material with iterations > 0 is active): can any
// pair of cells of the LDS region (the chunk and its HALO ring) take part in
// a rule under some random draw, can any cell change alone, or is any
// temperature stencil of the chunk non-zero?
template <int N>
__device__ bool tile_unstable(const MatTab &M, const uint32_t (*L)[LW], const uint32_t *T0, const int32_t *addt,
                              bool temp_enabled) {
    bool u = false;
    constexpr uint32_t KV = K_SINK | K_GRISE | K_REACT, KD = K_SINK | K_GRISE;
    // loops with a uniform counter (the wave's first index) and idle lanes
    // in the last round, rather than a loop under exec masks
    const int w0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~63u)), lane = threadIdx.x & 63;
    for (int i0 = w0; i0 < LW * LH; i0 += N) {
        const int i = i0 + lane;
        if (i >= LW * LH) continue;
        const int ly = i / LW, lx = i - ly * LW;
        const uint32_t c = L[ly][lx];
        const uint32_t t = cid(c) << 4;
        u |= single_unstable(M, c);
        if (lx + 1 < LW) u |= (T0[t | cid(L[ly][lx + 1])] & K_HSWAP) != 0;
        if (ly + 1 < LH) {
            u |= (T0[t | cid(L[ly + 1][lx])] & KV) != 0;
            if (lx + 1 < LW) u |= (T0[t | cid(L[ly + 1][lx + 1])] & KD) != 0;
            if (lx > 0) u |= (T0[t | cid(L[ly + 1][lx - 1])] & KD) != 0;
        }
    }
    if (temp_enabled) {
        for (int i0 = w0; i0 < TW * TH; i0 += N) {
            static_assert((TW * TH) % N == 0, "the chunk's cells fill whole rounds");
            const int i = i0 + lane;
            const int ly = HALO + i / TW, lx = HALO + (i & (TW - 1));
            const uint32_t c = L[ly][lx];
            if (cid(c) == WALL) continue;
            const int T = ctemp(c);
            const int v = temp_of(M, addt, T, cid(c), ntemp<true>(T, L[ly - 1][lx]), ntemp<true>(T, L[ly + 1][lx]),
                                  ntemp<true>(T, L[ly][lx - 1]), ntemp<true>(T, L[ly][lx + 1]));
            u |= v != T;
        }
    }
    return u;
}

// Blocks a wave handles in one phase: at most ceil(36 * 36 / N) rounds of 64,
// for a workgroup of N threads.
template <int N>
constexpr int qcap() { return ((36 * 36 + N - 1) / N) * 64; }

// One Margolus phase over the blocks whose cells the rest of the tick still
// reads: after this phase the tick needs `need` more rings around the chunk
// (need = cell_iter - it, so the region shrinks by one ring per phase and the
// last phase leaves the one ring the temperature stencil reads).
//
// Two passes per wave.  The first tests every block against the pair table
// and queues the ones some rule may change (ballot + mbcnt into the wave's
// LDS queue Q): "plain" blocks (no reaction, no gas or horizontal relation:
// sand over air, most of a world) from the front, the others from the back.
// The second pass runs the random draw and the rule on the queued blocks
// only, plain ones with the decision form that omits the gas and horizontal
// terms, so no lane idles through the decision for an inert block.  Blocks
// of a phase are disjoint and a wave's LDS accesses stay in order, so the
// passes need no barrier.  With TRACK, returns the band bits of chunk cells
// that changed.
// Block origins (LDS coordinates) whose block holds a tick-zone cell:
// [x0, x0 + w) x [y0, y0 + h); `on` is false when every block does.
struct ZoneBlocks {
    bool on;
    int x0, y0;
    uint32_t w, h;
};

// EXT: the side of the square of chunk cells the tick computes, from chunk
// coordinates (cx0, cy0): the whole chunk, or a quarter of it.
// DIRECT: no classification; each thread decides one block at a time with
// the general form (the same result for inert and plain blocks), for
// workgroups with about as many threads as the phase has blocks, where the
// classification pass only lengthens the chain of dependent steps.
template <int N, bool TRACK, bool REACT, bool ZONE, int EXT = TW, bool DIRECT = false>
__device__ __forceinline__ uint32_t run_phase(const MatTab &M, uint32_t (*L)[LW], const uint32_t *T, int it,
                                              uint64_t key, int ox, int oy, int need, int gx0, int gy0,
                                              bool has_wood, uint16_t *queue, const uint32_t *perm,
                                              const ZoneBlocks &zb, const uint32_t *ct, int cx0 = 0, int cy0 = 0) {
    static_assert(EXT == TW || !TRACK, "band tracking covers whole chunks");
    // first/last block origin (chunk coordinates) with the phase's parity
    const int xlo = cx0 - need, ylo = cy0 - need;
    const int x0f = xlo - ((xlo ^ ox) & 1), y0f = ylo - ((ylo ^ oy) & 1);
    const int x0l = (cx0 + EXT - 1 + need) - (((cx0 + EXT - 1 + need) ^ ox) & 1);
    const int y0l = (cy0 + EXT - 1 + need) - (((cy0 + EXT - 1 + need) ^ oy) & 1);
    const int nbx = ((x0l - x0f) >> 1) + 1, nby = ((y0l - y0f) >> 1) + 1;
    const int nb = nbx * nby;
    const uint32_t magic = (65536u + (uint32_t)nbx - 1u) / (uint32_t)nbx;  // b / nbx for b < 2^16/nbx
    const int lx0 = HALO + x0f, ly0 = HALO + y0f;
    // LDS coordinates of block b's top-left cell, and that cell
    auto block_xy = [&](int b, int &lx, int &ly) {
        const int by = (int)(__umul24((uint32_t)b, magic) >> 16);
        const int bx = b - (int)__umul24((uint32_t)by, (uint32_t)nbx);
        lx = lx0 + 2 * bx;
        ly = ly0 + 2 * by;
    };
    auto block_at = [&](int b) -> uint32_t * {
        int lx, ly;
        block_xy(b, lx, ly);
        return &L[0][0] + __umul24((uint32_t)ly, (uint32_t)LW) + lx;
    };
    auto in_zone = [&](int b) -> bool {
        const int by = (int)(__umul24((uint32_t)b, magic) >> 16);
        const int bx = b - (int)__umul24((uint32_t)by, (uint32_t)nbx);
        return (uint32_t)(lx0 + 2 * bx - zb.x0) < zb.w && (uint32_t)(ly0 + 2 * by - zb.y0) < zb.h;
    };
    const int lane = threadIdx.x & 63;
    constexpr int QCAP = qcap<N>();
    // the wave's first block as a uniform: the compiler cannot tell that
    // threadIdx.x & ~63 is one, and would run the loop under exec masks
    const int wb0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~63u));
    uint16_t *Q = queue + (wb0 >> 6) * QCAP;
    int nplain = 0, nother = 0;  // wave-uniform
    for (int b0 = wb0; b0 < nb && !DIRECT; b0 += N) {
        const int b = b0 + lane;
        // lanes past the last block test the last one again and queue nothing
        const uint32_t *p0 = block_at(b < nb ? b : nb - 1);
        const uint32_t q0 = p0[0], q1 = p0[1], q2 = p0[LW], q3 = p0[LW + 1];
        bool act, plain;
        if (ct) {  // one class-table read (ids < 8: at most 8 materials and no wall in the region)
            const uint32_t nib = (ct[((q0 & 7u) << 6) | ((q1 & 7u) << 3) | (q2 & 7u)] >> ((q3 & 7u) << 2)) & 15u;
            act = ((nib | ((nib >> 2) & (q0 >> 8)) | ((nib >> 3) & (q1 >> 8))) & 1u) != 0;
            plain = (nib & 2u) != 0;
        } else {
            const uint32_t i0 = q0 & 15u, i1 = q1 & 15u, i2 = q2 & 15u, i3 = q3 & 15u;
            const uint32_t VL = T[(i0 << 4) | i2], VR = T[(i1 << 4) | i3];
            const uint32_t DL = T[(i0 << 4) | i3], DR = T[(i1 << 4) | i2];
            const uint32_t HT = T[(i0 << 4) | i1], HB = T[(i2 << 4) | i3];
            act = !block_inert(VL, VR, DL, DR, HT, HB, (q0 >> 8) & 1u, (q1 >> 8) & 1u);
            plain = ((((VL | VR) & (K_GRISE | K_REACT)) | ((DL | DR) & K_GRISE) | ((HT | HB) & K_HSWAP)) == 0);
        }
        if (REACT && has_wood && (wood_hot(M, q0) || wood_hot(M, q1) || wood_hot(M, q2) || wood_hot(M, q3))) {
            act = true;
            plain = false;
        }
        act = act && b < nb;
        if (ZONE) act = act && in_zone(b);  // workgroups on the zone's edge only
        plain = plain && act;
        const uint64_t mp = __builtin_amdgcn_ballot_w64(plain), mo = __builtin_amdgcn_ballot_w64(act && !plain);
        const int bp = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mp >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mp, 0u));
        const int bo = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mo >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mo, 0u));
        if (act) Q[plain ? nplain + bp : QCAP - 1 - (nother + bo)] = (uint16_t)b;
        nplain += __popcll(mp);
        nother += __popcll(mo);
    }
    uint32_t bands = 0;
    auto track = [&](bool wrote, int lx, int ly) {
        if (TRACK && wrote) {
            const int cx = lx - HALO, cy = ly - HALO;  // chunk coordinates of the block
            if (cx + 1 >= 0 && cx <= TW - 1 && cy + 1 >= 0 && cy <= TH - 1)
                bands |= band_bits(max(cx, 0), min(cx + 1, TW - 1), max(cy, 0), min(cy + 1, TH - 1));
        }
    };
    // perm[y] holds, in byte k, the LDS offset (from the block's top-left
    // cell) of the destination of cell k under scatter mask y
    auto scatter = [&](uint32_t *p0, const uint32_t (&q)[4], uint32_t dec) {
        const uint32_t po = perm[dec & 0xFFu];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            // MOVED set on cell k (dec bit 8 + k), then cleared (bit 12 + k)
            const uint32_t v = (q[k] | ((dec >> k) & (uint32_t)MOVED)) & ~((dec >> (k + 4)) & (uint32_t)MOVED);
            p0[(po >> (8 * k)) & 0xFFu] = v;
        }
    };
    // uniform loops over the queues (rounds of 64 blocks)
    for (int qi = 0; qi < nplain; qi += 64) {
        const int i = qi + lane;
        if (i >= nplain) continue;  // the last round's idle lanes (not a divergent exit)
        int lx, ly;
        block_xy(Q[i], lx, ly);
        uint32_t *p0 = &L[0][0] + __umul24((uint32_t)ly, (uint32_t)LW) + lx;
        const uint32_t q[4] = {p0[0], p0[1], p0[LW], p0[LW + 1]};
        const uint32_t i0 = q[0] & 15u, i1 = q[1] & 15u, i2 = q[2] & 15u, i3 = q[3] & 15u;
        uint32_t r, s;
        block_hash(key, gx0 + lx, gy0 + ly, r, s);
        const uint32_t dec = block_decide<false>(T[(i0 << 4) | i2], T[(i1 << 4) | i3], T[(i0 << 4) | i3],
                                                 T[(i1 << 4) | i2], 0u, T[(i1 << 4) | i0], 0u,
                                                 (q[0] >> 8) & 1u, (q[1] >> 8) & 1u, 0u, 0u, r, s);
        if (dec != 0) scatter(p0, q, dec);
        track(dec != 0, lx, ly);
    }
    // the general form of a block: reactions, gas and horizontal relations
    auto general = [&](int b) {
        int lx, ly;
        block_xy(b, lx, ly);
        uint32_t *p0 = &L[0][0] + __umul24((uint32_t)ly, (uint32_t)LW) + lx;
        uint32_t q[4] = {p0[0], p0[1], p0[LW], p0[LW + 1]};
        const uint32_t i0 = q[0] & 15u, i1 = q[1] & 15u, i2 = q[2] & 15u, i3 = q[3] & 15u;
        const uint32_t VL = T[(i0 << 4) | i2], VR = T[(i1 << 4) | i3];
        bool hot = false;
        if (REACT && has_wood)
            hot = wood_hot(M, q[0]) || wood_hot(M, q[1]) || wood_hot(M, q[2]) || wood_hot(M, q[3]);
        // without classification: a wave whose blocks are all inert stops here
        if (DIRECT && !__builtin_amdgcn_ballot_w64(
                          hot || !block_inert(VL, VR, T[(i0 << 4) | i3], T[(i1 << 4) | i2], T[(i0 << 4) | i1],
                                              T[(i2 << 4) | i3], (q[0] >> 8) & 1u, (q[1] >> 8) & 1u)))
            return;
        uint32_t r, s;
        block_hash(key, gx0 + lx, gy0 + ly, r, s);
        bool wrote = false;
        if (REACT && (hot || ((VL | VR) & K_REACT))) {
            if (block_react(M, q)) {
                p0[0] = q[0];
                p0[1] = q[1];
                p0[LW] = q[2];
                p0[LW + 1] = q[3];
                wrote = true;
            }
        } else {
            const uint32_t dec = block_decide<true>(VL, VR, T[(i0 << 4) | i3], T[(i1 << 4) | i2], T[(i0 << 4) | i1],
                                                    T[(i1 << 4) | i0], T[(i2 << 4) | i3], (q[0] >> 8) & 1u,
                                                    (q[1] >> 8) & 1u, (q[2] >> 8) & 1u, (q[3] >> 8) & 1u, r, s);
            if (dec != 0) scatter(p0, q, dec);
            wrote = dec != 0;
        }
        track(wrote, lx, ly);
    };
    if (DIRECT) {
        for (int b0 = wb0; b0 < nb; b0 += N) {
            const int b = b0 + lane;
            if (b >= nb) continue;
            if (ZONE && !in_zone(b)) continue;
            general(b);
        }
        return bands;
    }
    for (int qi = 0; qi < nother; qi += 64) {
        const int i = qi + lane;
        if (i >= nother) continue;
        general(Q[QCAP - 1 - i]);
    }
    return bands;
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
    return v;
}

// Final compact words of a thread's 2 x ROWS output cells (columns lx, lx + 1,
// rows ly0 .. ly0 + ROWS - 1), with the temperature step applied when temp_now; the
// stencil slides down the strip.  With TCHG, returns whether a temperature
// changed (only chunk skipping needs to know).  With ZONE, only cells of the
// tick zone take the step: LDS rows [zr0, zr1), the thread's columns whose
// bits (1: lx, 2: lx + 1) are set in zc.
template <int ROWS, bool WALLS, bool TCHG = true, bool ZONE = false>
__device__ __forceinline__ bool final_words(const MatTab &M, const uint32_t (*L)[LW], const int32_t *addt,
                                            bool temp_now, int lx, int ly0, uint32_t (&fw)[2 * ROWS],
                                            int zr0 = 0, int zr1 = 0, uint32_t zc = 0) {
    bool tchg = false;
    if (!temp_now) {
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const uint2 c = *reinterpret_cast<const uint2 *>(&L[ly0 + r][lx]);
            fw[2 * r] = c.x;
            fw[2 * r + 1] = c.y;
        }
        return false;
    }
    if (!WALLS && !M.any_add_temp) {
        // T + ((Tu + Td + Tl + Tr - 4T) >> 3) is T when the four neighbours
        // share T, and no material adds heat: a wave whose strips and their
        // neighbours all hold one temperature keeps its words as they are.
        const uint32_t w0 = L[ly0][lx];
        uint32_t acc = 0;  // OR of the differences from w0; temperatures in bits 16-31
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const uint2 c = *reinterpret_cast<const uint2 *>(&L[ly0 + r][lx]);
            fw[2 * r] = c.x;
            fw[2 * r + 1] = c.y;
            acc |= (c.x ^ w0) | (c.y ^ w0) | (L[ly0 + r][lx - 1] ^ w0) | (L[ly0 + r][lx + 2] ^ w0);
        }
        const uint2 up = *reinterpret_cast<const uint2 *>(&L[ly0 - 1][lx]);
        const uint2 dn = *reinterpret_cast<const uint2 *>(&L[ly0 + ROWS][lx]);
        acc |= (up.x ^ w0) | (up.y ^ w0) | (dn.x ^ w0) | (dn.y ^ w0);
        if (__all((acc >> 16) == 0)) return false;
    }
    uint2 up = *reinterpret_cast<const uint2 *>(&L[ly0 - 1][lx]);
    uint2 cur = *reinterpret_cast<const uint2 *>(&L[ly0][lx]);
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const uint2 dn = *reinterpret_cast<const uint2 *>(&L[ly0 + r + 1][lx]);
        const uint32_t cl = L[ly0 + r][lx - 1], cr = L[ly0 + r][lx + 2];
        const int t0 = ctemp(cur.x), t1 = ctemp(cur.y);
        int v0 = t0, v1 = t1;
        if (!WALLS || cid(cur.x) != WALL)
            v0 = temp_of(M, addt, t0, cid(cur.x), ntemp<WALLS>(t0, up.x), ntemp<WALLS>(t0, dn.x),
                         ntemp<WALLS>(t0, cl), ntemp<WALLS>(t0, cur.y));
        if (!WALLS || cid(cur.y) != WALL)
            v1 = temp_of(M, addt, t1, cid(cur.y), ntemp<WALLS>(t1, up.y), ntemp<WALLS>(t1, dn.y),
                         ntemp<WALLS>(t1, cur.x), ntemp<WALLS>(t1, cr));
        if (ZONE) {
            const bool rin = ly0 + r >= zr0 && ly0 + r < zr1;
            if (!(rin && (zc & 1u))) v0 = t0;
            if (!(rin && (zc & 2u))) v1 = t1;
        }
        if (TCHG) tchg |= (v0 != t0) | (v1 != t1);
        fw[2 * r] = (cur.x & 0xFFFFu) | ((uint32_t)v0 << 16);
        fw[2 * r + 1] = (cur.y & 0xFFFFu) | ((uint32_t)v1 << 16);
        up = cur;
        cur = dn;
    }
    return tchg;
}

// Region map of the load and of the colour plane.  The region is 72 rows of
// 36 two-cell pairs.  At 256 threads, thread t < 252 owns pair column t % 36 of
// the rows 1 + t / 36 + 7k (k < 10), which are rows 1 .. 70: every row a cell of the
// chunk can come from in one tick (MAX_IT rows at most), so these are also
// the colours the thread holds.  Threads t < 72 load the two edge rows 0
// and 71 besides.  The rows of one thread are 7 apart, so its addresses
// differ by constants: the address arithmetic is done once, not per pair.
// (For N = 512 threads: 14 rows a sweep, 5 sweeps; for 1024: 28 rows a sweep,
// and the third sweep covers the last 14 rows only.)
constexpr int PR = LW / 2;                         // two-cell pairs per region row
template <int N> constexpr int rows_per_sweep() { return N / PR; }                // 7 at 256 threads
template <int N> constexpr int sweeps() {                                         // 10 at 256 threads
    return (LH - 2 + rows_per_sweep<N>() - 1) / rows_per_sweep<N>();
}
template <int N> constexpr bool whole_sweeps() { return rows_per_sweep<N>() * sweeps<N>() == LH - 2; }
static_assert(whole_sweeps<256>() && whole_sweeps<512>() && HALO - MAX_IT >= 1, "sweeps cover rows 1 .. LH-2 exactly");
static_assert((uint64_t)FSG_MAX_SIDE * 8u * LH < (1ull << 32), "32-bit byte offsets within a region");

// The tick zone against a workgroup's region [gx0, gx0 + LW) x [gy0, gy0 + LH):
// dead when no cell of the region lies in the zone (then no block the
// workgroup computes acts and no chunk cell takes the temperature step);
// blocks.on when the region's in-world part is not wholly inside the zone.
struct ZoneTest {
    bool dead;
    ZoneBlocks blocks;
};
__device__ __forceinline__ ZoneTest zone_test(const DevWorld &w, int gx0, int gy0) {
    ZoneTest z;
    z.dead = gx0 + LW <= w.zx0 || gx0 >= w.zx1 || gy0 + LH <= w.zy0 || gy0 >= w.zy1;
    const bool inside = max(gx0, 0) >= w.zx0 && min(gx0 + LW, w.W) <= w.zx1 && max(gy0, 0) >= w.zy0 &&
                        min(gy0 + LH, w.H) <= w.zy1;
    z.blocks.on = !z.dead && !inside;
    // a block at LDS origin (lx, ly) holds a zone cell iff
    // zx0 - 1 <= gx0 + lx < zx1 and zy0 - 1 <= gy0 + ly < zy1
    z.blocks.x0 = w.zx0 - 1 - gx0;
    z.blocks.y0 = w.zy0 - 1 - gy0;
    z.blocks.w = (uint32_t)(w.zx1 - w.zx0 + 1);
    z.blocks.h = (uint32_t)(w.zy1 - w.zy0 + 1);
    return z;
}

// Wake test of the skip path: the chunk may change by itself (F_POT), changed
// last tick (F_ANY: its two state buffers differ), a neighbour changed inside
// the band the chunk's region overlaps, or the tick is forced.
__device__ __forceinline__ bool chunk_awake(const DevWorld &w, const uint32_t *fr, int tx, int ty, bool forced) {
    auto fl = [&](int dx, int dy) -> uint32_t {
        const int nx = tx + dx, ny = ty + dy;
        return (nx < 0 || ny < 0 || nx >= w.tiles_x || ny >= w.tiles_y) ? 0u : fr[ny * w.tiles_x + nx];
    };
    return forced || (fl(0, 0) & (F_POT | F_ANY)) || (fl(0, -1) & F_S) || (fl(0, 1) & F_N) || (fl(-1, 0) & F_E) ||
           (fl(1, 0) & F_W) || (fl(-1, -1) & F_SE) || (fl(1, -1) & F_SW) || (fl(-1, 1) & F_NE) ||
           (fl(1, 1) & F_NW);
}

// SK / RE: chunk skipping / reactions compiled in (1), out (0), or read
// from the world at run time (2, the batch kernels, whose worlds may differ).
// SK == 3: skipping on, and the chunk is known to be awake (it was taken from
// the awake list k_wake built).  ZN: the tick zone is tested (1), or the
// world's zone covers it and no zone code is compiled in (0).
template <int SK, int RE, int ZN, int N = NT>
__device__ __forceinline__ void tile_update(const DevWorld &w, uint64_t tick, int tile, uint32_t (*L)[LW],
                                            uint32_t (*lut)[256], uint32_t *rgba, int32_t *addt,
                                            uint16_t *queue, uint32_t *red) {
    static_assert(N == 256 || N == 512 || N == 1024, "the load map and the write-back strips");
    constexpr int RS = rows_per_sweep<N>(), KC = sweeps<N>();
    // sweep k of this thread's rows lies in rows 1 .. LH-2 (always, unless the
    // last sweep is partial)
    auto row_in = [&](int k) { return whole_sweeps<N>() || (int)(threadIdx.x / PR) + RS * k < LH - 2; };
    constexpr int ROWS = TH * 32 / N;  // rows of a thread's write-back strip (8 at 256 threads)
    const int tid = threadIdx.x;
    const MatTab &M = w.mat;
    const bool react = RE == 2 ? M.reactions != 0 : RE == 1;
    const int tx = tile % w.tiles_x, ty = tile / w.tiles_x;
    const int gx0 = tx * TW - HALO, gy0 = ty * TH - HALO;  // world coords of L[0][0]
    const uint64_t *__restrict__ in = w.buf[tick & 1];
    uint64_t *__restrict__ out = w.buf[(tick + 1) & 1];
    const bool skip = SK == 2 ? w.skip != 0 : SK != 0;
    bool prev_changed = true;
    // Tick zone: no block of the region holds a zone cell (the chunk cannot
    // change), or some do and some do not (blocks are tested one by one).
    ZoneTest zt{};
    if (ZN) zt = zone_test(w, gx0, gy0);

    if (skip) {
        // Exact skip: a chunk sleeps unless it may change by itself (F_POT)
        // or a neighbour changed inside the band its region overlaps.
        const uint32_t *fr = w.flags[(tick + 1) & 1];
        const uint32_t self = fr[tile];
        const bool forced = tick == w.force_tick;
        if (SK != 3 && !chunk_awake(w, fr, tx, ty, forced)) {
            if (tid == 0) w.flags[tick & 1][tile] = 0;
            return;
        }
        // both state buffers hold this chunk unless it changed last tick
        prev_changed = forced || (self & F_ANY);
        if (zt.dead && !prev_changed) {
            if (tid == 0) w.flags[tick & 1][tile] = 0;
            return;
        }
    }

    // Every global load (pair table, colour and add_temp tables, the region
    // as 16-byte cell pairs) is issued before the first LDS write, so a
    // workgroup waits for one memory latency, not one per loop trip.
    const int ci = (int)M.cell_iter;
    uint32_t pv[MAX_IT];
    const bool tab = N == 256 || tid < 256;  // one pair-table entry per thread
#pragma unroll
    for (int k = 0; k < MAX_IT; k++)
        if (k < ci && tab) pv[k] = w.pair[k][tid];
    uint32_t mrgba = 0;
    int32_t madd = 0;
    if (tid < 16) {
        mrgba = M.rgba[tid];
        madd = M.add_temp[tid];
    }
    // A chunk whose halo lies inside an even-width world takes the fast
    // paths: 16-byte loads/stores with no bounds or wall checks.
    const bool interior = gx0 >= 0 && gy0 >= 0 && gx0 + LW <= w.W && gy0 + LH <= w.H && (w.W & 1) == 0;
    // the classification reads the world's class tables when it has them and
    // the region holds no wall (its ids are then all below 8); the slot in
    // LDS holds the table of the phase's iteration.  Not in the 1024-thread
    // form: a small world's tick is one chunk's latency, and the refill's
    // barrier between phases cost it 3% (512^2 world, 12.8 -> 13.2 ms for
    // 1000 ticks)
    const bool use_ct = N == NT && w.ct != nullptr && interior;
    uint2 ctv = make_uint2(0u, 0u);
    if (use_ct && tab) ctv = reinterpret_cast<const uint2 *>(w.ct)[tid];
    const bool mine = tid < RS * PR, edge = tid < 2 * PR;
    const int rr = tid / PR, c2 = tid - rr * PR;                      // sweep rows 1 + rr + RS * k
    const int ey = tid < PR ? 0 : LH - 1, ec2 = tid < PR ? tid : tid - PR;  // edge row (edge threads)
    uint4 v[KC], ve;
    if (interior) {
        // a uniform row base per sweep plus one 32-bit offset per thread
        const char *__restrict__ src = reinterpret_cast<const char *>(in + (size_t)gy0 * w.W + gx0);
        const uint32_t rowb = (uint32_t)w.W * 8u;
        const uint32_t off = (uint32_t)(1 + rr) * rowb + 16u * (uint32_t)c2;
        if (mine) {
#pragma unroll
            for (int k = 0; k < KC; k++)
                if (row_in(k)) v[k] = *reinterpret_cast<const uint4 *>(src + (off + (uint32_t)(RS * k) * rowb));
        }
        if (edge) ve = *reinterpret_cast<const uint4 *>(src + (uint32_t)ey * rowb + 16u * (uint32_t)ec2);
    } else {
        auto load = [&](int ly, int p) -> uint4 {
            const int gx = gx0 + 2 * p, gy = gy0 + ly;
            const bool rin = gy >= 0 && gy < w.H;
            const int64_t row = (int64_t)gy * w.W;
            const uint64_t a = (rin && gx >= 0 && gx < w.W) ? in[row + gx] : WALL;
            const uint64_t b = (rin && gx + 1 >= 0 && gx + 1 < w.W) ? in[row + gx + 1] : WALL;
            return make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
        };
        if (mine) {
#pragma unroll
            for (int k = 0; k < KC; k++)
                if (row_in(k)) v[k] = load(1 + rr + RS * k, c2);
        }
        if (edge) ve = load(ey, ec2);
    }
    // compact cells into LDS (origin residues in bits 9-14), colours kept in
    // registers until the write-back
    uint32_t cola[KC], colb[KC];  // colours of the pair's two cells
    bool wood = false;
    if (mine) {
        uint32_t *Lt = &L[1 + rr][2 * c2];
        const uint32_t orgx = ((uint32_t)(2 * c2) & 7u) << 9, oy = (uint32_t)(1 + rr) << 12;
#pragma unroll
        for (int k = 0; k < KC; k++) {
            if (!row_in(k)) continue;
            const uint32_t org = orgx | ((oy + ((uint32_t)(RS * k) << 12)) & 0x7000u);  // (row & 7) << 12
            *reinterpret_cast<uint2 *>(Lt + RS * k * LW) = make_uint2(v[k].x | org, v[k].z | org | (1u << 9));
            cola[k] = v[k].y;
            colb[k] = v[k].w;
        }
    }
    if (edge) {
        const uint32_t org = (((uint32_t)(2 * ec2) & 7u) << 9) | (((uint32_t)ey & 7u) << 12);
        *reinterpret_cast<uint2 *>(&L[ey][2 * ec2]) = make_uint2(ve.x | org, ve.z | org | (1u << 9));
    }
    if (RE != 0 && react) {
        if (mine) {
#pragma unroll
            for (int k = 0; k < KC; k++)
                if (row_in(k)) wood |= (v[k].x & 0xFFu) == M.wood || (v[k].z & 0xFFu) == M.wood;
        }
        if (edge) wood |= (ve.x & 0xFFu) == M.wood || (ve.z & 0xFFu) == M.wood;
    }
#pragma unroll
    for (int k = 0; k < MAX_IT; k++)
        if (k < ci && tab) lut[k][tid] = pv[k];
    if (tab) {
        // scatter offsets for mask y = tid: cell k goes to k ^ ((y >> 2k) & 3)
        uint32_t po = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t d = (uint32_t)k ^ (((uint32_t)tid >> (2 * k)) & 3u);
            po |= ((d & 1u) + (d >> 1) * (uint32_t)LW) << (8 * k);
        }
        lut[MAX_IT][tid] = po;
        if (use_ct) reinterpret_cast<uint2 *>(lut[MAX_IT + 1])[tid] = ctv;
    }
    if (tid < 16) {
        rgba[tid] = mrgba;
        addt[tid] = madd;
    }
    // burning wood is the one reaction the pair table cannot see (the
    // reduction is also the barrier after the staging)
    bool has_wood = false;
    if (RE == 0) __syncthreads();
    else has_wood = __syncthreads_or(wood && react) != 0;

    if (skip && !prev_changed) {
        // the chunk did not change last tick: if nothing in its region can
        // change now it keeps its state, and both buffers already hold it
        // (the test ignores the tick zone: it may only wake too many)
        const bool pot = __syncthreads_or(tile_unstable<N>(M, L, lut[0], addt, M.temp_every != 0)) != 0;
        if (!pot) {
            if (tid == 0) {
                w.flags[tick & 1][tile] = 0;
                atomicAdd(w.processed, 1ull);
            }
            return;
        }
    }

    uint32_t bands = 0;
    if (!zt.dead) {
        const uint64_t seedmix = splitmix64(w.seed);
        const uint32_t *ct = use_ct ? lut[MAX_IT + 1] : nullptr;
        for (int it = 0; it < ci; it++) {
            const uint64_t g = tick * (uint64_t)M.cell_iter + (uint64_t)it;
            const uint64_t key = splitmix64(seedmix ^ g);
            const int ox = kOX[g & 3], oy = kOY[g & 3];
            // the next iteration's class table, loaded while this phase runs
            const bool next_ct = use_ct && it + 1 < ci;
            if (next_ct && tab) ctv = reinterpret_cast<const uint2 *>(w.ct + (it + 1) * 512)[tid];
            if (zt.blocks.on)  // on the zone's edge: one form for every setting (the bands may go unused)
                bands |= run_phase<N, true, RE != 0, true>(M, L, lut[it], it, key, ox, oy, ci - it, gx0, gy0, has_wood,
                                                        queue, lut[MAX_IT], zt.blocks, ct);
            else if (SK != 0 && skip)
                bands |= run_phase<N, true, RE != 0, false>(M, L, lut[it], it, key, ox, oy, ci - it, gx0, gy0, has_wood,
                                                         queue, lut[MAX_IT], zt.blocks, ct);
            else if (SK != 1 && SK != 3)
                run_phase<N, false, RE != 0, false>(M, L, lut[it], it, key, ox, oy, ci - it, gx0, gy0, has_wood, queue,
                                                 lut[MAX_IT], zt.blocks, ct);
            __syncthreads();
            if (next_ct) {  // every wave's classification of this phase is done
                if (tab) reinterpret_cast<uint2 *>(lut[MAX_IT + 1])[tid] = ctv;
                __syncthreads();
            }
        }
    }

    // Write-back.  Thread = 2-column strip x ROWS rows of the chunk.
    const bool temp_now = !zt.dead && M.temp_every != 0 && (tick % M.temp_every) == M.temp_phase;
    const int j = tid & 31, rg = tid >> 5;
    const int lx = HALO + 2 * j, ly0 = HALO + rg * ROWS;
    uint32_t fw[2 * ROWS];
    bool tchg;
    if (zt.blocks.on) {
        // on the zone's edge: cells outside it keep their temperature (the
        // wall-checking form serves interior chunks too)
        const int gx = gx0 + lx;
        const uint32_t zc = (gx >= w.zx0 && gx < w.zx1 ? 1u : 0u) | (gx + 1 >= w.zx0 && gx + 1 < w.zx1 ? 2u : 0u);
        tchg = final_words<ROWS, true, true, true>(M, L, addt, temp_now, lx, ly0, fw, w.zy0 - gy0, w.zy1 - gy0, zc);
    } else {
        tchg = skip ? (interior ? final_words<ROWS, false>(M, L, addt, temp_now, lx, ly0, fw)
                                : final_words<ROWS, true>(M, L, addt, temp_now, lx, ly0, fw))
                    : (interior ? final_words<ROWS, false, false>(M, L, addt, temp_now, lx, ly0, fw)
                                : final_words<ROWS, true, false>(M, L, addt, temp_now, lx, ly0, fw));
    }
    if (skip && tchg) bands |= band_bits(2 * j, 2 * j + 1, rg * ROWS, rg * ROWS + ROWS - 1);
    __syncthreads();
    // the colour plane: each cell's colour at its origin (rows 1 .. LH-2)
    if (mine) {
        uint32_t *Lt = &L[1 + rr][2 * c2];
#pragma unroll
        for (int k = 0; k < KC; k++) {
            if (!row_in(k)) continue;
            Lt[RS * k * LW] = cola[k];
            Lt[RS * k * LW + 1] = colb[k];
        }
    }
    __syncthreads();
    // A reaction product takes its material's base colour; any other cell
    // the colour of its origin, which is
    // within 3 cells in x and y: the signed 3-bit differences of the origin
    // residues and the cell's own coordinates.
    // (The offset dy * LW + dx is pinned as two v_bfe_i32 and one
    // v_mad_i32_i24: left to itself the compiler rewrites the extracts as
    // shift pairs and folds y into a 32-bit multiply, ~10 instead of 6 VALU
    // per cell.  Register-only instructions.)
    const uint32_t *Lf = &L[0][0];
    auto origin = [&](uint32_t wd, int x, int y) -> int {
        int dx, dy, d;
        asm("v_bfe_i32 %0, %1, 9, 3" : "=v"(dx) : "v"(wd - ((uint32_t)x << 9)));    // bits 9-11
        asm("v_bfe_i32 %0, %1, 12, 3" : "=v"(dy) : "v"(wd - ((uint32_t)y << 12)));  // bits 12-14
        asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(dy), "s"(LW), "v"(dx));
        return d + (y * LW + x);
    };
    auto colour = [&](uint32_t wd, int x, int y) -> uint32_t {
        return *((wd & C_XF) ? rgba + (wd & 15u) : Lf + origin(wd, x, y));
    };
    const bool xf_on = RE != 0 && react;  // only reactions make XF cells
    if (interior) {
        // a uniform base plus a 32-bit offset per row
        char *__restrict__ dstb = reinterpret_cast<char *>(out + (size_t)(gy0 + HALO) * w.W + (gx0 + HALO));
        const uint32_t rowb = (uint32_t)w.W * 8u;
        const uint32_t voff = (uint32_t)(ly0 - HALO) * rowb + 16u * (uint32_t)j;
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const uint32_t a = fw[2 * r], b = fw[2 * r + 1];
            uint32_t ca, cb;
            if (xf_on) {
                ca = colour(a, lx, ly0 + r);
                cb = colour(b, lx + 1, ly0 + r);
            } else {
                ca = Lf[origin(a, lx, ly0 + r)];
                cb = Lf[origin(b, lx + 1, ly0 + r)];
            }
            *reinterpret_cast<uint4 *>(dstb + (voff + (uint32_t)r * rowb)) = make_uint4(a & C_KEEP, ca, b & C_KEEP, cb);
        }
    } else {
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const int gy = gy0 + ly0 + r;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int gx = gx0 + lx + c;
                if (gx < w.W && gy < w.H) {
                    const uint32_t a = fw[2 * r + c];
                    out[(size_t)gy * w.W + gx] = (uint64_t)(a & C_KEEP) | ((uint64_t)colour(a, lx + c, ly0 + r) << 32);
                }
            }
        }
    }
    if (skip) {
        bands = wave_or(bands);
        if ((tid & 63) == 0) red[tid >> 6] = bands;
        __syncthreads();
        if (tid == 0) {
            uint32_t all = 0;
            for (int k = 0; k < N / 64; k++) all |= red[k];
            w.flags[tick & 1][tile] = (zt.dead ? 0u : F_POT) | all;  // a dead chunk cannot change by itself
            atomicAdd(w.processed, 1ull);
        }
    }
}

// Position in a row-major order of n chunks taken by workgroup i.  The
// hardware deals workgroups to the eight XCDs in turn (i % 8), each with its
// own L2.  In every run of 8g positions, XCD k takes the g consecutive ones
// from kg on, so horizontally adjacent chunks, which share the 128-byte lines
// of their halo columns, run back to back on one L2, while the grid still
// sweeps the world in order (a tail shorter than 8g keeps its order).  With
// g = tx / 8 for a world of tx chunks a row, the 16384^2 world fetches 1.01x
// its algorithmic bytes instead of 1.26x.
__device__ __forceinline__ int xcd_order(int i, int n, int g) {
    const int run = 8 * g, full = n - n % run;
    if (i >= full) return i;
    const int b = i / run, r = i - b * run;  // r = 8 * (slot in XCD) + XCD
    return b * run + (r & 7) * g + (r >> 3);
}

// Small worlds (at most one chunk per compute unit): the same update with
// SMALL_NT (1024) threads per chunk.  A small world leaves compute units idle,
// so a tick lasts one workgroup's latency; four times the waves per chunk cut
// each wave's share of the phases and let four waves per SIMD hide each
// other's LDS latency.
template <int RE, int ZN>
__global__ __launch_bounds__(SMALL_NT) void k_tick_small(DevWorld w, uint32_t k) {
    __shared__ __align__(16) uint32_t Lmem[LH * LW + 16];  // the region, then the 16 base colours
    __shared__ uint32_t lut[MAX_IT + 3][256];  // pair tables, scatter offsets, the phase's class table
    uint32_t(*L)[LW] = reinterpret_cast<uint32_t(*)[LW]>(Lmem);
    uint32_t *rgba = Lmem + LH * LW;
    __shared__ int32_t addt[16];
    __shared__ uint16_t queue[(SMALL_NT / 64) * qcap<SMALL_NT>()];
    __shared__ uint32_t red[SMALL_NT / 64];
    tile_update<0, RE, ZN, SMALL_NT>(w, w.tick + k, blockIdx.x, L, lut, rgba, addt, queue, red);
}
// Smaller worlds still (at most a quarter chunk per compute unit): each
// chunk as four 32x32 quarters, one workgroup of QUARTER_NT threads each,
// with the HALO ring around the quarter.  Such a tick lasts one workgroup's
// chain of dependent steps (stage, the phases, write-back); a quarter's
// window is (32 + 2 HALO)^2 cells instead of (64 + 2 HALO)^2, so at 1024
// threads each thread decides one block per phase, with no classification
// pass (512^2 world: 8.9 us a tick instead of 11.2 with one 1024-thread
// workgroup per chunk).  The same update, at 256 threads and classified,
// runs the last round's chunks of k_tick_world.  The window sits at its
// place in the chunk's LDS layout, so block origins, hash positions, origin
// residues and the phase schedule are the chunk update's, and the quarter's
// cells come out the same: every cell of the quarter depends, over the
// tick's phases, only on cells of its window (one ring per phase plus the
// temperature stencil's).  No chunk skipping here.  ZN: the tick zone is
// tested, with the chunk's zone test (a window inside a dead region is dead;
// on the zone's edge every block and cell is tested).
constexpr int QW = TW / 2, QH = TH / 2;                   // the quarter (32 x 32)
constexpr int QLW = QW + 2 * HALO, QLH = QH + 2 * HALO;   // its window (40 x 40)
static_assert(TW == 2 * QW && TH == 2 * QH && (QW & 1) == 0, "quarters of whole cell pairs");
// SEP: the colours have an LDS array of their own (colw, QLW x QLH, the
// window's cells), written at the stage, so the write-back needs no colour
// plane and no barrier; otherwise they go through registers and the
// region array after the phases (the chunk update's way, which keeps LDS
// for occupancy).
template <int RE, int ZN, int N, bool SEP = false>
__device__ __forceinline__ void quarter_update(const DevWorld &w, uint64_t tick, int tile, int quad,
                                               uint32_t (*L)[LW], uint32_t (*lut)[256], uint32_t *rgba,
                                               int32_t *addt, uint16_t *queue, uint32_t *colw = nullptr) {
    static_assert(N == 256 || N == 512 || N == 1024, "one pair-table entry per thread, write-back strips");
    const int tid = threadIdx.x;
    const MatTab &M = w.mat;
    const int tx = tile % w.tiles_x, ty = tile / w.tiles_x;
    const int gx0 = tx * TW - HALO, gy0 = ty * TH - HALO;    // world coords of L[0][0], as for the chunk
    const int cx0 = (quad & 1) * QW, cy0 = (quad >> 1) * QH;  // the quarter (chunk coords) = its window (LDS coords)
    ZoneTest zt{};
    if (ZN) zt = zone_test(w, gx0, gy0);
    const uint64_t *__restrict__ in = w.buf[tick & 1];
    uint64_t *__restrict__ out = w.buf[(tick + 1) & 1];
    const int ci = (int)M.cell_iter;
    uint32_t pv[MAX_IT];
#pragma unroll
    for (int k = 0; k < MAX_IT; k++)
        if (k < ci && tid < 256) pv[k] = w.pair[k][tid];
    uint32_t mrgba = 0;
    int32_t madd = 0;
    if (tid < 16) {
        mrgba = M.rgba[tid];
        madd = M.add_temp[tid];
    }
    // the window: QLH rows of QP cell pairs, world corner (wx0, wy0)
    constexpr int QP = QLW / 2, NPAIR = QLH * QP, KQ = (NPAIR + N - 1) / N;
    const int wx0 = gx0 + cx0, wy0 = gy0 + cy0;
    const bool interior = wx0 >= 0 && wy0 >= 0 && wx0 + QLW <= w.W && wy0 + QLH <= w.H && (w.W & 1) == 0;
    uint4 v[KQ];
#pragma unroll
    for (int k = 0; k < KQ; k++) {
        const int i = tid + k * N;
        if (i >= NPAIR) continue;
        const int row = i / QP, p = i - row * QP;
        const int gx = wx0 + 2 * p, gy = wy0 + row;
        if (interior) {
            v[k] = *reinterpret_cast<const uint4 *>(in + (size_t)gy * w.W + gx);
        } else {
            const bool rin = gy >= 0 && gy < w.H;
            const int64_t rb = (int64_t)gy * w.W;
            const uint64_t a = (rin && gx >= 0 && gx < w.W) ? in[rb + gx] : WALL;
            const uint64_t b = (rin && gx + 1 >= 0 && gx + 1 < w.W) ? in[rb + gx + 1] : WALL;
            v[k] = make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
        }
    }
    // compact cells with their origin residues (LDS coordinates mod 8); the
    // colours go to colw (SEP) or stay in registers until the write-back
    uint32_t cola[KQ], colb[KQ];
    bool wood = false;
#pragma unroll
    for (int k = 0; k < KQ; k++) {
        const int i = tid + k * N;
        if (i >= NPAIR) continue;
        const int row = i / QP, p = i - row * QP;
        const int lx = cx0 + 2 * p, ly = cy0 + row;
        const uint32_t org = (((uint32_t)lx & 7u) << 9) | (((uint32_t)ly & 7u) << 12);
        *reinterpret_cast<uint2 *>(&L[ly][lx]) = make_uint2(v[k].x | org, v[k].z | org | (1u << 9));
        if (SEP) {
            *reinterpret_cast<uint2 *>(&colw[row * QLW + 2 * p]) = make_uint2(v[k].y, v[k].w);
        } else {
            cola[k] = v[k].y;
            colb[k] = v[k].w;
        }
        if (RE != 0) wood |= (v[k].x & 0xFFu) == M.wood || (v[k].z & 0xFFu) == M.wood;
    }
#pragma unroll
    for (int k = 0; k < MAX_IT; k++)
        if (k < ci && tid < 256) lut[k][tid] = pv[k];
    if (tid < 256) {
        // scatter offsets for mask y = tid: cell k goes to k ^ ((y >> 2k) & 3)
        uint32_t po = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t d = (uint32_t)k ^ (((uint32_t)tid >> (2 * k)) & 3u);
            po |= ((d & 1u) + (d >> 1) * (uint32_t)LW) << (8 * k);
        }
        lut[MAX_IT][tid] = po;
    }
    if (tid < 16) {
        rgba[tid] = mrgba;
        addt[tid] = madd;
    }
    bool has_wood = false;
    if (RE == 0) __syncthreads();
    else has_wood = __syncthreads_or(wood) != 0;

    const uint64_t seedmix = splitmix64(w.seed);
    for (int it = 0; it < ci && !zt.dead; it++) {
        const uint64_t g = tick * (uint64_t)M.cell_iter + (uint64_t)it;
        const uint64_t key = splitmix64(seedmix ^ g);
        // at 1024 threads a thread per block (at most 20 x 20 a phase): no classification
        constexpr bool DIRECT = N >= 1024;
        if (ZN && zt.blocks.on)
            run_phase<N, false, RE != 0, true, QW, DIRECT>(M, L, lut[it], it, key, kOX[g & 3], kOY[g & 3], ci - it,
                                                           gx0, gy0, has_wood, queue, lut[MAX_IT], zt.blocks, nullptr,
                                                           cx0, cy0);
        else
            run_phase<N, false, RE != 0, false, QW, DIRECT>(M, L, lut[it], it, key, kOX[g & 3], kOY[g & 3], ci - it,
                                                            gx0, gy0, has_wood, queue, lut[MAX_IT], zt.blocks, nullptr,
                                                            cx0, cy0);
        __syncthreads();
    }

    // Write-back.  Thread = 2-column strip x QROWS rows of the quarter (the
    // first 512 threads, whole waves, at 1024).
    constexpr int QSTRIPS = QH * (QW / 2);                      // 512 pair strips of one row
    constexpr int QROWS = QSTRIPS > N ? QSTRIPS / N : 1;        // 2 at 256 threads, else 1
    const bool writer = tid < QSTRIPS / QROWS;
    const bool temp_now = !zt.dead && M.temp_every != 0 && (tick % M.temp_every) == M.temp_phase;
    const int j = tid & (QW / 2 - 1), rg = tid / (QW / 2);
    const int lx = HALO + cx0 + 2 * j, ly0 = HALO + cy0 + rg * QROWS;
    uint32_t fw[2 * QROWS];
    if (writer) {
        if (ZN && zt.blocks.on) {
            // on the zone's edge: cells outside it keep their temperature
            const int gx = gx0 + lx;
            const uint32_t zc = (gx >= w.zx0 && gx < w.zx1 ? 1u : 0u) | (gx + 1 >= w.zx0 && gx + 1 < w.zx1 ? 2u : 0u);
            final_words<QROWS, true, false, true>(M, L, addt, temp_now, lx, ly0, fw, w.zy0 - gy0, w.zy1 - gy0, zc);
        } else if (interior) {
            final_words<QROWS, false, false>(M, L, addt, temp_now, lx, ly0, fw);
        } else {
            final_words<QROWS, true, false>(M, L, addt, temp_now, lx, ly0, fw);
        }
    }
    if (!SEP) {
        __syncthreads();
        // the colour plane: each cell's colour at its origin
#pragma unroll
        for (int k = 0; k < KQ; k++) {
            const int i = tid + k * N;
            if (i >= NPAIR) continue;
            const int row = i / QP, p = i - row * QP;
            L[cy0 + row][cx0 + 2 * p] = cola[k];
            L[cy0 + row][cx0 + 2 * p + 1] = colb[k];
        }
        __syncthreads();
    }
    const uint32_t *Lf = &L[0][0];
    auto colour = [&](uint32_t wd, int x, int y) -> uint32_t {
        if (RE != 0 && (wd & C_XF)) return rgba[wd & 15u];
        const int dx = (int)((((wd >> 9) - (uint32_t)x) & 7u) ^ 4u) - 4;   // signed 3-bit differences
        const int dy = (int)((((wd >> 12) - (uint32_t)y) & 7u) ^ 4u) - 4;
        if (SEP) return colw[(y + dy - cy0) * QLW + (x + dx - cx0)];
        return Lf[(y + dy) * LW + x + dx];
    };
#pragma unroll
    for (int r = 0; r < QROWS; r++) {
        if (!writer) break;
        const int ly = ly0 + r, gy = gy0 + ly;
        const uint32_t a = fw[2 * r], b = fw[2 * r + 1];
        const uint32_t ca = colour(a, lx, ly), cb = colour(b, lx + 1, ly);
        const int gx = gx0 + lx;
        if (interior) {
            *reinterpret_cast<uint4 *>(out + (size_t)gy * w.W + gx) = make_uint4(a & C_KEEP, ca, b & C_KEEP, cb);
        } else {
            if (gy < w.H && gx < w.W) out[(size_t)gy * w.W + gx] = (uint64_t)(a & C_KEEP) | ((uint64_t)ca << 32);
            if (gy < w.H && gx + 1 < w.W) out[(size_t)gy * w.W + gx + 1] = (uint64_t)(b & C_KEEP) | ((uint64_t)cb << 32);
        }
    }
}

template <int RE, int ZN>
__global__ __launch_bounds__(QUARTER_NT) void k_tick_quarter(DevWorld w, uint32_t k) {
    __shared__ __align__(16) uint32_t Lmem[LH * LW + 16];  // the chunk's region layout, then the 16 base colours
    __shared__ uint32_t lut[MAX_IT + 3][256];  // pair tables, scatter offsets
    uint32_t(*L)[LW] = reinterpret_cast<uint32_t(*)[LW]>(Lmem);
    uint32_t *rgba = Lmem + LH * LW;
    __shared__ int32_t addt[16];
    __shared__ uint16_t queue[(QUARTER_NT / 64) * qcap<QUARTER_NT>()];
    __shared__ __align__(8) uint32_t colw[QLW * QLH];  // the window's colours
    quarter_update<RE, ZN, QUARTER_NT, true>(w, w.tick + k, (int)(blockIdx.x >> 2), (int)(blockIdx.x & 3u), L, lut,
                                             rgba, addt, queue, colw);
}
template __global__ void k_tick_quarter<0, 0>(DevWorld, uint32_t);
template __global__ void k_tick_quarter<1, 0>(DevWorld, uint32_t);
template __global__ void k_tick_quarter<0, 1>(DevWorld, uint32_t);
template __global__ void k_tick_quarter<1, 1>(DevWorld, uint32_t);

template __global__ void k_tick_small<0, 0>(DevWorld, uint32_t);
template __global__ void k_tick_small<1, 0>(DevWorld, uint32_t);
template __global__ void k_tick_small<0, 1>(DevWorld, uint32_t);
template __global__ void k_tick_small<1, 1>(DevWorld, uint32_t);

// Five workgroups per CU (at most 96 VGPRs, 28 KB of LDS each).  One
// instantiation per (chunk skipping, reactions) setting of the world, so
// the branches and paths a setting never takes are not compiled in.
// Workgroups [0, nfull) update whole chunks in the XCD-aware order; the
// chunks from nfull on are updated as quarters (four workgroups each, the
// grid's last), so the launch's last round of workgroups is short: a chunk
// left over from the last full round of whole chunks would hold one
// compute unit for a whole chunk's latency while the others idle.
template <int SK, int RE, int ZN>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_tick_world(DevWorld w, uint32_t k,
                                                                                            int nfull) {
    __shared__ __align__(16) uint32_t Lmem[LH * LW + 16];  // the region, then the 16 base colours
    __shared__ uint32_t lut[MAX_IT + 3][256];  // pair tables, scatter offsets, the phase's class table
    uint32_t(*L)[LW] = reinterpret_cast<uint32_t(*)[LW]>(Lmem);
    uint32_t *rgba = Lmem + LH * LW;
    __shared__ int32_t addt[16];
    __shared__ uint16_t queue[(NT / 64) * qcap<NT>()];
    __shared__ uint32_t red[NT / 64];
    const int i = (int)blockIdx.x;
    if (i < nfull) {
        tile_update<SK, RE, ZN>(w, w.tick + k, xcd_order(i, nfull, (w.tiles_x & 7) ? 8 : w.tiles_x >> 3), L, lut,
                                rgba, addt, queue, red);
    } else {
        // the four quarters of a chunk on one XCD (workgroup i runs on XCD
        // i % 8, and nfull is a multiple of 8), so they share their halo
        // lines in one L2: in each run of 32 workgroups, those of one XCD
        // take the four quarters of one chunk (a tail shorter than 32 keeps
        // the plain order)
        const int q = i - nfull, nq = (int)gridDim.x - nfull, full = nq & ~31;
        const int c = q < full ? (q & 7) + ((q >> 5) << 3) : q >> 2;
        const int quad = q < full ? (q >> 3) & 3 : q & 3;
        quarter_update<RE, ZN, NT>(w, w.tick + k, nfull + c, quad, L, lut, rgba, addt, queue);
    }
}
template __global__ void k_tick_world<0, 0, 0>(DevWorld, uint32_t, int);
template __global__ void k_tick_world<0, 1, 0>(DevWorld, uint32_t, int);
template __global__ void k_tick_world<0, 0, 1>(DevWorld, uint32_t, int);
template __global__ void k_tick_world<0, 1, 1>(DevWorld, uint32_t, int);

// The awake chunks of tick t = w.tick + k, listed for k_tick_list: one thread
// per chunk runs the wake test of the skip path (and writes the flags of the
// chunks that sleep), then each workgroup reserves its slice of the list
// with one atomic on the count awake_n[t & 1], which is zero on entry: this
// launch clears awake_n[(t + 1) & 1] for the next tick (tick t - 1, the
// last to read it, has finished).  The list's order is that of the
// reservations; the result does not depend on it (the chunks of one tick
// are independent).
__global__ __launch_bounds__(256) void k_wake(DevWorld w, uint32_t k) {
    __shared__ uint32_t wave_n[4], base;
    const uint64_t tick = w.tick + k;
    const int ntiles = w.tiles_x * w.tiles_y;
    const int tile = (int)(blockIdx.x * 256 + threadIdx.x);
    if (tile == 0) w.awake_n[(tick + 1) & 1] = 0;
    bool awake = false;
    if (tile < ntiles) {
        const int tx = tile % w.tiles_x, ty = tile / w.tiles_x;
        const uint32_t *fr = w.flags[(tick + 1) & 1];
        const bool forced = tick == w.force_tick;
        awake = chunk_awake(w, fr, tx, ty, forced);
        // a chunk outside the zone's reach only needs a tick to copy itself
        if (awake && !forced && !(fr[tile] & F_ANY) && zone_test(w, tx * TW - HALO, ty * TH - HALO).dead)
            awake = false;
        if (!awake) w.flags[tick & 1][tile] = 0;
    }
    const uint64_t m = __ballot(awake);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) wave_n[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t n = wave_n[0] + wave_n[1] + wave_n[2] + wave_n[3];
        base = n ? atomicAdd(&w.awake_n[tick & 1], n) : 0u;
    }
    __syncthreads();
    if (awake) {
        uint32_t off = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        for (int v = 0; v < wv; v++) off += wave_n[v];
        w.awake[off] = (uint32_t)tile;
    }
}

// Skip-path tick over the awake list: workgroup i updates the i-th listed
// chunk, and workgroups past the list's end exit at once.  The listed
// chunks take the first workgroups, so they start back to back and no
// sleeping chunk holds a workgroup slot while they run.
template <int RE, int ZN>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_tick_list(DevWorld w, uint32_t k) {
    __shared__ __align__(16) uint32_t Lmem[LH * LW + 16];  // the region, then the 16 base colours
    __shared__ uint32_t lut[MAX_IT + 3][256];  // pair tables, scatter offsets, the phase's class table
    uint32_t(*L)[LW] = reinterpret_cast<uint32_t(*)[LW]>(Lmem);
    uint32_t *rgba = Lmem + LH * LW;
    __shared__ int32_t addt[16];
    __shared__ uint16_t queue[(NT / 64) * qcap<NT>()];
    __shared__ uint32_t red[NT / 64];
    const uint64_t tick = w.tick + k;
    const int n = (int)w.awake_n[tick & 1];
    if ((int)blockIdx.x >= n) return;
    // the list runs in row-major order within each reserving workgroup of k_wake
    tile_update<3, RE, ZN>(w, tick, (int)w.awake[xcd_order(blockIdx.x, n, 8)], L, lut, rgba, addt, queue, red);
}
template __global__ void k_tick_list<0, 0>(DevWorld, uint32_t);
template __global__ void k_tick_list<1, 0>(DevWorld, uint32_t);
template __global__ void k_tick_list<0, 1>(DevWorld, uint32_t);
template __global__ void k_tick_list<1, 1>(DevWorld, uint32_t);

// Many worlds in one launch: tile_start[i] is the first global chunk index of
// world i (prefix sums, tile_start[n] = total).
__device__ __forceinline__ void tick_batch_chunk(const DevWorld *__restrict__ worlds,
                                                 const uint32_t *__restrict__ tile_start, int nworlds, uint32_t k) {
    __shared__ __align__(16) uint32_t Lmem[LH * LW + 16];  // the region, then the 16 base colours
    __shared__ uint32_t lut[MAX_IT + 3][256];  // pair tables, scatter offsets, the phase's class table
    uint32_t(*L)[LW] = reinterpret_cast<uint32_t(*)[LW]>(Lmem);
    uint32_t *rgba = Lmem + LH * LW;
    __shared__ int32_t addt[16];
    __shared__ uint16_t queue[(NT / 64) * qcap<NT>()];
    __shared__ uint32_t red[NT / 64];
    const uint32_t gt = blockIdx.x;
    int lo = 0, hi = nworlds - 1;
    while (lo < hi) {  // last world with tile_start <= gt
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= gt) lo = mid;
        else hi = mid - 1;
    }
    const DevWorld &w = worlds[lo];
    tile_update<2, 2, 1>(w, w.tick + k, (int)(gt - tile_start[lo]), L, lut, rgba, addt, queue, red);
}

// Two register budgets of the batch tick, one of the autotuner's choices:
// five workgroups per CU (95 VGPRs) or four (up to 128).
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_tick_batch(
    const DevWorld *__restrict__ worlds, const uint32_t *__restrict__ tile_start, int nworlds, uint32_t k) {
    tick_batch_chunk(worlds, tile_start, nworlds, k);
}
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_tick_batch4(
    const DevWorld *__restrict__ worlds, const uint32_t *__restrict__ tile_start, int nworlds, uint32_t k) {
    tick_batch_chunk(worlds, tile_start, nworlds, k);
}

__global__ void k_advance(DevWorld *worlds, int nworlds, uint32_t ticks) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nworlds) worlds[i].tick += ticks;
}

// Standalone temperature step (world::tickTemperature): reads buf[t&1], writes
// buf[(t+1)&1]; the caller advances the buffer parity by swapping roles.
__global__ __launch_bounds__(NT) void k_temperature(const uint64_t *__restrict__ in,
                                                     uint64_t *__restrict__ out, int W, int H,
                                                     MatTab M, int4 zone) {
    const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    if (i >= (size_t)W * H) return;
    const int x = (int)(i % (size_t)W), y = (int)(i / (size_t)W);
    const uint64_t c = in[i];
    if (x < zone.x || y < zone.y || x >= zone.z || y >= zone.w) {  // outside the tick zone
        out[i] = c;
        return;
    }
    const int T = ctemp(c);
    const int tu = y > 0 ? ctemp(in[i - W]) : T;
    const int td = y < H - 1 ? ctemp(in[i + W]) : T;
    const int tl = x > 0 ? ctemp(in[i - 1]) : T;
    const int tr = x < W - 1 ? ctemp(in[i + 1]) : T;
    int v = T + ((tu + td + tl + tr - 4 * T) >> 3);
    if (M.any_add_temp) v += M.add_temp[cid(c)];
    v = v > 32767 ? 32767 : (v < -32768 ? -32768 : v);
    out[i] = (c & 0xFFFFFFFF0000FFFFull) | ((uint64_t)(uint16_t)v << 16);
}

// Census: per-material counts and an order-sensitive checksum
// (sum of splitmix64(cell + index*GOLDEN)); out[0..15] counts, out[16] checksum.
__global__ __launch_bounds__(NT) void k_stats(const uint64_t *__restrict__ cells, size_t n,
                                              unsigned long long *out) {
    __shared__ unsigned int hist[16];
    __shared__ unsigned long long wsum[NT / 64];
    if (threadIdx.x < 16) hist[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long sum = 0;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        const uint64_t c = cells[i];
        atomicAdd(&hist[c & 15u], 1u);
        sum += splitmix64(c + (uint64_t)i * GOLDEN);
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x < 16 && hist[threadIdx.x]) atomicAdd(&out[threadIdx.x], (unsigned long long)hist[threadIdx.x]);
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < NT / 64; k++) t += wsum[k];
        atomicAdd(&out[16], t);
    }
}

// setTile batch (edits deduplicated on the host, so order does not matter).
__global__ void k_fill(uint64_t *cells, size_t n, uint64_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cells[i] = v;
}

// Streaming copy, one 16-byte load and store per thread over a grid that
// covers the buffer: the calibration kernel for the HBM byte counters and
// the reference for attainable HBM bandwidth.
__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                              size_t n16) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}

__global__ void k_set_cells(uint64_t *cells, int W, int H, const int32_t *xy, const uint64_t *vals,
                            size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = xy[2 * i], y = xy[2 * i + 1];
    if (x < 0 || y < 0 || x >= W || y >= H) return;  // setTile ignores OOB (world.cpp:1005)
    cells[(size_t)y * W + x] = vals[i];
}

// ---------------------------------------------------------------------------
// Rigid-body stamp / unstamp (game.cpp:1760-1815, 1896-1983) in 5 claim
// rounds.  A claim is the 64-bit value (round << 32 | pixel); atomicMin on
// the target's slot in the spare state buffer makes the lowest pixel of the
// earliest round win, which the CPU mirrors by scanning pixels in order.
// ---------------------------------------------------------------------------
__constant__ int kDX[5] = {0, 1, -1, 0, 0};
__constant__ int kDY[5] = {0, 0, 0, 1, -1};

// The pixel of this thread (blockIdx.y = world, x over its pixels) and, for
// round d (d = -1: the pixel's own cell), its target cell.  The own cell of a
// live pixel (not air, body present and enabled) is computed once per stamp
// or unstamp by k_claims_init into tgt; a round adds its offset.
struct PixelRef {
    const StampWorld *w;
    uint32_t p;
};

__device__ __forceinline__ bool pixel_ref(const StampArgs &a, PixelRef &ref) {
    ref.w = &a.worlds[blockIdx.y];
    ref.p = blockIdx.x * blockDim.x + threadIdx.x;
    return ref.p < ref.w->npix;
}

constexpr uint32_t NOT_LIVE = 0x80000000u;

__device__ __forceinline__ bool pixel_target(const PixelRef &ref, int d, int &wx, int &wy) {
    const uint64_t t = ref.w->tgt[ref.p];
    if ((uint32_t)t == NOT_LIVE) return false;
    wx = (int)(uint32_t)t;
    wy = (int)(uint32_t)(t >> 32);
    if (d >= 0) {
        wx += kDX[d];
        wy += kDY[d];
    }
    return wx >= 0 && wy >= 0 && wx < ref.w->W && wy < ref.w->H;
}

// Q16.16 value to int, rounding toward zero like the reference's (int) cast.
__device__ __forceinline__ int q16_trunc(int64_t v) { return (int)(v >= 0 ? v >> 16 : -((-v) >> 16)); }

// Own cell of pixel p under its body's pose (Q16.16 rotation of the pixel's
// lower-left corner, as the reference rotates (tx, ty + 1), truncated toward
// zero: game.cpp:1774-1775), or NOT_LIVE.
__device__ __forceinline__ uint64_t own_cell(const StampArgs &a, const StampWorld &W, uint32_t p) {
    if (bit(a.mat.air, cid(W.pixels[p]))) return NOT_LIVE;
    const uint32_t b = W.pix_body[p];
    if (b == 0xFFFFFFFFu) return NOT_LIVE;
    const fsg_body B = W.bodies[b];
    if (!B.enabled) return NOT_LIVE;
    const uint32_t k = p - B.pixel_offset;
    const int tx = (int)(k % (uint32_t)B.mat_w), ty = (int)(k / (uint32_t)B.mat_w);
    const int64_t fx = (int64_t)tx * B.cos_q16 - (int64_t)(ty + 1) * B.sin_q16 + B.pos_x_q16;
    const int64_t fy = (int64_t)tx * B.sin_q16 + (int64_t)(ty + 1) * B.cos_q16 + B.pos_y_q16;
    return (uint64_t)(uint32_t)q16_trunc(fx) | ((uint64_t)(uint32_t)q16_trunc(fy) << 32);
}

// Chunks whose spare-buffer cells serve as claim slots must run on the next
// tick (the slots were written); flags is null without skipping.  Takes the
// wave's lanes together (flags is the same in every lane): one atomic per
// distinct chunk of the wave, not one per lane -- a wave's pixels fall in one
// or two chunks, and same-address atomics serialise.
__device__ __forceinline__ void mark_tiles_wave(uint32_t *flags, bool on, uint32_t tile) {
    const int lane = threadIdx.x & 63;
    uint64_t m = __ballot(on);
    while (m) {
        const int leader = __builtin_ctzll(m);
        const uint32_t t0 = (uint32_t)__shfl((int)tile, leader, 64);
        if (lane == leader) atomicOr(&flags[t0], F_WAKE_ALL);
        on = on && tile != t0;
        m = __ballot(on);
    }
}

// Every live pixel's own cell, and the claim slots of its five targets (the
// own cell and its four neighbours) cleared before any round.  A slot that
// receives proposals in some round gets a winner, which fills the cell
// (stamp) or empties it (unstamp); a later round's proposal to that cell
// then either fails the cell check of its propose step or, for a SAND or
// SOUP typed body pixel (which leaves the cell a stamping target), loses
// the atomicMin to the earlier round's claim still in the slot -- so the
// earliest round and then the lowest pixel win, as a round-by-round pass
// over the pixels gives.  With save the slots' contents (the state before
// the last tick) are first copied out, and k_claims_clear clears them in a
// launch of its own, after every copy; k_claims_restore puts them back.
__device__ __forceinline__ void slot_at(const StampWorld &W, uint64_t t, int d, bool live, bool &on, int &wx,
                                        int &wy) {
    wx = (int)(uint32_t)t + kDX[d];
    wy = (int)(uint32_t)(t >> 32) + kDY[d];
    on = live && wx >= 0 && wy >= 0 && wx < W.W && wy < W.H;
}

__global__ void k_claims_init(StampArgs a, int save) {
    const StampWorld &W = a.worlds[blockIdx.y];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = p < W.npix;
    uint64_t t = NOT_LIVE;
    if (in) {
        W.state[p] = 0;
        t = own_cell(a, W, p);
        W.tgt[p] = t;
    }
    const bool live = in && (uint32_t)t != NOT_LIVE;
    for (int d = 0; d < 5; d++) {
        bool on;
        int wx, wy;
        slot_at(W, t, d, live, on, wx, wy);
        if (save) {
            if (on) W.save[5 * (size_t)p + d] = W.claims[(size_t)wy * W.W + wx];
            continue;
        }
        if (on) W.claims[(size_t)wy * W.W + wx] = ~0ull;
        if (W.flags) mark_tiles_wave(W.flags, on, on ? (uint32_t)((wy / TH) * W.tiles_x + wx / TW) : 0u);
    }
}

__global__ void k_claims_clear(StampArgs a) {
    const StampWorld &W = a.worlds[blockIdx.y];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = p < W.npix;
    const uint64_t t = in ? W.tgt[p] : NOT_LIVE;
    const bool live = in && (uint32_t)t != NOT_LIVE;
    for (int d = 0; d < 5; d++) {
        bool on;
        int wx, wy;
        slot_at(W, t, d, live, on, wx, wy);
        if (on) W.claims[(size_t)wy * W.W + wx] = ~0ull;
        if (W.flags) mark_tiles_wave(W.flags, on, on ? (uint32_t)((wy / TH) * W.tiles_x + wx / TW) : 0u);
    }
}

// Every pixel that saved a slot writes the saved value back; pixels sharing
// a slot saved the same value, so the writes agree.
__global__ void k_claims_restore(StampArgs a) {
    const StampWorld &W = a.worlds[blockIdx.y];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= W.npix || !W.save) return;
    const uint64_t t = W.tgt[p];
    const bool live = (uint32_t)t != NOT_LIVE;
    for (int d = 0; d < 5; d++) {
        bool on;
        int wx, wy;
        slot_at(W, t, d, live, on, wx, wy);
        if (on) W.claims[(size_t)wy * W.W + wx] = W.save[5 * (size_t)p + d];
    }
}

// The per-pixel steps of a claim round (d: round).  Each runs in two launch
// forms: over all pixels (grid: y = world, x = pixel), and for rounds after
// the first over the compacted list of pixels still open after round 0
// (most pixels are placed in round 0, so later rounds touch only the few
// left instead of reading every pixel's state).

__device__ __forceinline__ void stamp_step(const StampArgs &a, const PixelRef &ref, int op, int d) {
    const StampWorld &W = *ref.w;
    if (W.state[ref.p]) return;
    int wx, wy;
    if (!pixel_target(ref, op == OP_UNFINISH ? -1 : d, wx, wy)) return;
    const size_t ci = (size_t)wy * W.W + wx;
    const uint64_t mine = ((uint64_t)d << 32) | ref.p;
    switch (op) {
    case OP_PROPOSE: {
        const uint32_t m = cid(W.cells[ci]);
        if (bit(a.mat.air, m) || bit(a.mat.sand, m) || bit(a.mat.soup, m))
            atomicMin(reinterpret_cast<unsigned long long *>(&W.claims[ci]), (unsigned long long)mine);
        break;
    }
    case OP_RESOLVE: {
        if (W.claims[ci] != mine) break;
        const uint64_t old = W.cells[ci];
        if (!bit(a.mat.air, cid(old))) {
            W.disp[ref.p] = old;
            W.disp_xy[ref.p] = (uint64_t)(uint32_t)wx | ((uint64_t)(uint32_t)wy << 32);
            W.state[ref.p] = bit(a.mat.sand, cid(old)) ? 2 : 3;  // placed over sand / soup
        } else {
            W.state[ref.p] = 1;  // placed
        }
        W.cells[ci] = W.pixels[ref.p];
        break;
    }
    case OP_UNPROPOSE:
        if (cbyte(W.cells[ci]) == cbyte(W.pixels[ref.p]))
            atomicMin(reinterpret_cast<unsigned long long *>(&W.claims[ci]), (unsigned long long)mine);
        break;
    case OP_UNRESOLVE: {
        if (W.claims[ci] != mine) break;
        const uint64_t c = W.cells[ci];
        W.pixels[ref.p] = c;
        W.cells[ci] = (uint64_t)a.mat.air_id | (c & 0xFFFF0000ull);
        W.state[ref.p] = 1;
        break;
    }
    case OP_UNFINISH:  // a pixel that found nothing is lost when the cell under it is air
        if (bit(a.mat.air, cid(W.cells[ci]))) W.pixels[ref.p] = (uint64_t)a.mat.air_id;
        break;
    }
}

__global__ void k_stamp_step(StampArgs a, int op, int d) {
    PixelRef ref;
    if (!pixel_ref(a, ref)) return;
    stamp_step(a, ref, op, d);
}

__global__ void k_stamp_step_list(StampArgs a, int op, int d) {
    const uint32_t n = *a.nopen;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t e = a.open[i];
        PixelRef ref;
        ref.w = &a.worlds[(uint32_t)(e >> 32)];
        ref.p = (uint32_t)e;
        stamp_step(a, ref, op, d);
    }
}

// The live pixels still open after round 0, appended to a.open.  Nearly
// every wave of a rotated body holds a pixel that lost its round-0 claim, so
// the list is reserved once per workgroup of STAMP_CPT * 256 pixels (one
// global atomic per workgroup, offsets by a workgroup prefix sum) rather than
// once per wave, whose same-address atomics serialise.
__global__ __launch_bounds__(256) void k_stamp_compact(StampArgs a) {
    __shared__ uint32_t wtot[4], base_s;
    const StampWorld &W = a.worlds[blockIdx.y];
    const uint32_t p0 = blockIdx.x * (256u * STAMP_CPT) + threadIdx.x;
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < STAMP_CPT; k++) {
        const uint32_t p = p0 + 256u * k;
        if (p < W.npix && !W.state[p] && (uint32_t)W.tgt[p] != NOT_LIVE) bits |= 1u << k;
    }
    const uint32_t mine = __popc(bits);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = mine;  // inclusive prefix over the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wtot[wv] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        base_s = t ? atomicAdd(a.nopen, t) : 0u;
    }
    __syncthreads();
    if (!bits) return;
    uint32_t k = base_s + incl - mine;
    for (int w = 0; w < wv; w++) k += wtot[w];
    const uint64_t hi = (uint64_t)blockIdx.y << 32;
#pragma unroll
    for (int b = 0; b < STAMP_CPT; b++)
        if (bits >> b & 1u) a.open[k++] = hi | (p0 + 256u * b);
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0; }

} // namespace fsg
