// fsg_device.hpp -- device-side data layout and the per-block update rule of
// the deterministic falling-sand tick, written for gfx950 (CDNA4).
//
// Material properties live in 64-bit nibble tables and 32-bit bit masks so a
// lookup is one shift and one mask on values that sit in SGPRs: no dependent
// loads through a Material* per neighbour (the reference chases tile.mat->...
// for every neighbour, world.cpp:1097-1211).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsg {

constexpr int TW = 64;               // chunk (tile) width in cells
constexpr int TH = 64;               // chunk height
constexpr int HALO = 4;              // cell_iter (<= 3) Margolus phases + 1 temperature step
constexpr int LW = TW + 2 * HALO;    // LDS tile width  (72)
constexpr int LH = TH + 2 * HALO;    // LDS tile height (72)
constexpr int NT = 256;              // threads per workgroup (4 wave64)
constexpr int SMALL_NT = 1024;       // threads per chunk of a world with at most one chunk per CU
constexpr int QUARTER_NT = 1024;     // threads per quarter chunk of a world with at most a quarter chunk per CU
constexpr int MAX_IT = 3;            // largest cell_iter the fused kernel supports
constexpr uint64_t MOVED = 0x100;
constexpr uint64_t WALL = 15;        // out-of-world cell: inert solid (getTile OOB, world.cpp:1000)
constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

struct MatTab {
    uint64_t pass;          // nibble/id: SOLID -> 15, else density rank
    uint64_t sink[MAX_IT];  // nibble/id: SAND or SOUP active in iteration it -> rank, else 0
    uint64_t flow[MAX_IT];  // nibble/id: SOUP active in it -> rank, else 0
    uint32_t gas[MAX_IT];   // bit/id: GAS active in it
    uint32_t air, sand, soup; // bit/id
    uint32_t rgba[16];      // colour of cells created by reactions
    int32_t add_temp[16];
    uint32_t any_add_temp;
    int32_t burn;           // wood burn temperature
    uint32_t water, lava, steam, stone, wood, air_id;
    uint32_t reactions, cell_iter, temp_every, temp_phase;
};

struct DevWorld {
    uint64_t *buf[2];       // state before tick t lives in buf[t & 1]
    uint32_t *flags[2];     // per-chunk activity; tick t reads flags[(t+1)&1], writes flags[t&1]
    int32_t W, H, tiles_x, tiles_y;
    uint64_t seed;
    uint64_t tick;          // tick of launch offset k = 0
    uint64_t force_tick;    // every chunk is processed at this tick
    uint32_t skip;          // exact dirty-rect skipping on
    uint32_t pad;
    unsigned long long *processed; // chunk updates run (skip mode)
    // tick zone: cells [zx0, zx1) x [zy0, zy1) tick; a block acts iff one of
    // its cells lies inside, and only zone cells take the temperature step
    int32_t zx0, zy0, zx1, zy1;
    // awake-chunk list of the skip path: awake[0 .. awake_n[t & 1]) are the
    // chunks tick t updates (k_wake, k_tick_list)
    uint32_t *awake;
    uint32_t *awake_n;
    // block-class tables of the classification pass, MAX_IT x 512 words
    // (worlds of at most 8 materials; null otherwise): word i0<<6|i1<<3|i2 of
    // table it holds, in nibble i3, the class of a block of material ids
    // i0 i1 / i2 i3 at iteration it: bit 0 some pair relation (K_SINK,
    // K_GRISE, K_HSWAP, K_REACT), bit 1 plain, bits 2-3 K_TSANDACT of the top
    // left / right cell
    const uint32_t *ct;
    MatTab mat;
    uint16_t pair[MAX_IT][256];  // pair_code(mat, it, t, u) at [it][t * 16 + u], built on the host
};

// Chunk activity flags (one u32 per chunk per tick parity):
//   F_POT   the chunk may change next tick even if no neighbour changes;
//   F_ANY   the chunk changed this tick;
//   F_N..F_SE  it changed within HALO cells of that edge / corner, the band
//           a neighbour's region (chunk + HALO ring) overlaps.
constexpr uint32_t F_POT = 1u, F_ANY = 2u;
constexpr uint32_t F_N = 4u, F_S = 8u, F_W = 16u, F_E = 32u;
constexpr uint32_t F_NW = 64u, F_NE = 128u, F_SW = 256u, F_SE = 512u;
constexpr uint32_t F_WAKE_ALL = 0x3FFu;  // changed everywhere: wakes the chunk and its 8 neighbours

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + GOLDEN;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__host__ __device__ __forceinline__ int nib(uint64_t tab, uint32_t m) { return (int)((tab >> (m << 2)) & 15u); }
__host__ __device__ __forceinline__ bool bit(uint32_t mask, uint32_t m) { return (mask >> m) & 1u; }
__host__ __device__ __forceinline__ uint32_t cid(uint64_t c) { return (uint32_t)c & 15u; }
__host__ __device__ __forceinline__ uint32_t cbyte(uint64_t c) { return (uint32_t)c & 0xFFu; }
__host__ __device__ __forceinline__ int ctemp(uint64_t c) { return (int)(int16_t)(uint16_t)(c >> 16); }

// Compact cell, the form the tick kernel keeps in LDS: the stored cell's low
// word, whose bits 9-15 are zero in every stored cell, with those bits reused:
//   bits 0-7   material byte, bit 8 MOVED              (as stored)
//   bits 9-11  origin x mod 8, bits 12-14 origin y mod 8 (LDS coordinates of
//              the cell at the start of the tick; a cell moves at most one
//              step per phase, so with |displacement| <= 3 the residues name
//              the origin, and moves stay plain permutations)
//   bit 15     XF: made by a reaction, so its colour is the material's base
//              colour rather than the origin cell's
//   bits 16-31 temperature                             (as stored)
// The colour (high word) stays outside LDS until the chunk is written back.
constexpr uint32_t C_ORG = 0x7E00u, C_XF = 0x8000u, C_KEEP = 0xFFFF01FFu;

// Reaction product: material m, same temperature, base colour, flags clear.
__host__ __device__ __forceinline__ uint64_t xform(const MatTab &M, uint64_t c, uint32_t m) {
    return (uint64_t)m | (c & 0xFFFF0000ull) | ((uint64_t)M.rgba[m & 15] << 32);
}
// The same on a compact cell: the colour is implied by XF.
__host__ __device__ __forceinline__ uint32_t xform(const MatTab &, uint32_t c, uint32_t m) {
    return m | (c & (0xFFFF0000u | C_ORG)) | C_XF;
}

// Block random words for the block whose top-left world cell is (x0, y0).
__host__ __device__ __forceinline__ void block_hash(uint64_t key, int x0, int y0, uint32_t &r, uint32_t &s) {
    const uint32_t pos = (uint32_t)(x0 + 1) | ((uint32_t)(y0 + 1) << 16);
    r = lowbias32(pos ^ (uint32_t)key);
    s = lowbias32(r ^ (uint32_t)(key >> 32));
}

// Stage R of the block rule, reactions: water directly above lava becomes
// steam and every lava cell of the block becomes stone; hot wood becomes
// steam and trades places with the block's first air cell (world.cpp:1181-1204,
// 1519-1537).  Returns true if a reaction fired (the block then does not also
// move).  Requires M.reactions.
template <typename Cell>
__host__ __device__ __forceinline__ bool block_react(const MatTab &M, Cell q[4]) {
    bool fired = false;
    const bool wl = cbyte(q[0]) == M.water && cbyte(q[2]) == M.lava;
    const bool wr = cbyte(q[1]) == M.water && cbyte(q[3]) == M.lava;
    if (wl || wr) {
        if (wl) q[0] = xform(M, q[0], M.steam);
        if (wr) q[1] = xform(M, q[1], M.steam);
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (cbyte(q[i]) == M.lava) q[i] = xform(M, q[i], M.stone);
        fired = true;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (cbyte(q[i]) == M.wood && ctemp(q[i]) > M.burn) {
            q[i] = xform(M, q[i], M.steam);
            bool done = false;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (!done && bit(M.air, cid(q[j]))) {
                    const Cell t = q[i];
                    q[i] = q[j];
                    q[j] = t;
                    done = true;
                }
            }
            fired = true;
        }
    }
    return fired;
}

// One Margolus block, q = {TL, TR, BL, BR} (stored cells, or compact cells
// in the kernel).  Returns true if a cell changed.  Stages: R reactions, V
// vertical, D diagonal, F friction stop, H horizontal; each stage restates
// one group of reference rules (world.cpp:1148-1889).
template <typename Cell>
__host__ __device__ __forceinline__ bool block_rule(const MatTab &M, int it, uint32_t r, uint32_t s, Cell q[4]) {
    const Cell o0 = q[0], o1 = q[1], o2 = q[2], o3 = q[3];
    const uint64_t pass = M.pass, sinkt = M.sink[it], flowt = M.flow[it];
    const uint32_t gasm = M.gas[it], airm = M.air, sandm = M.sand;

    if (M.reactions && block_react(M, q)) return q[0] != o0 || q[1] != o1 || q[2] != o2 || q[3] != o3;

    // V: per column, left then right
    bool vm = false;
#pragma unroll
    for (int col = 0; col < 2; col++) {
        const int o = 1 - col;
        Cell t = q[col];
        const Cell u = q[col + 2];
        const uint32_t mt = cid(t), mu = cid(u);
        const int sr = nib(sinkt, mt);
        if (nib(pass, mu) < sr) {
            const bool sand = bit(sandm, mt);
            const bool skip = sand && nib(pass, cid(q[o + 2])) < sr && ((r >> (4 * col)) & 15u) == 0;
            if (!skip) {
                if (sand && ((r >> (8 + col)) & 1u)) t |= MOVED;
                q[col] = u;
                q[col + 2] = t;
                vm = true;
                if (sand && ((r >> (10 + 2 * col)) & 3u) == 0 && bit(sandm, cid(q[o + 2])))
                    q[o + 2] |= MOVED;
            }
        } else if (bit(gasm, mu) && bit(airm, mt)) {
            if (!(bit(airm, cid(q[o])) && ((r >> (14 + col)) & 1u))) {
                q[col] = u;
                q[col + 2] = t;
                vm = true;
            }
        }
    }
    if (vm) return true;

    // D: TL->BR or TR->BL for sinking cells (sand needs MOVED, or the 1/16
    // unstick draw r[16:19] / r[20:23] == 0, and shouldMove), else gas BL->TR
    // or BR->TL into air
    {
        const uint32_t ma = cid(q[0]), mb = cid(q[1]);
        bool cl = nib(pass, cid(q[3])) < nib(sinkt, ma);
        bool cr = nib(pass, cid(q[2])) < nib(sinkt, mb);
        const bool ua = ((r >> 16) & 15u) == 0, ub = ((r >> 20) & 15u) == 0;
        if (cl && bit(sandm, ma)) {
            if (!(q[0] & MOVED) && !ua) cl = false;
            else if ((s & 31u) == 0) { cl = false; q[0] &= ~MOVED; }
        }
        if (cr && bit(sandm, mb)) {
            if (!(q[1] & MOVED) && !ub) cr = false;
            else if (((s >> 5) & 31u) == 0) { cr = false; q[1] &= ~MOVED; }
        }
        if (cl && cr) {
            if ((s >> 10) & 1u) cl = false;
            else cr = false;
        }
        if (cl) {
            Cell a2 = q[0];
            const bool sand = bit(sandm, ma);
            if (sand && ua) a2 |= MOVED;
            if (sand && ((s >> 11) & 511u) == 0) a2 &= ~MOVED;
            if (sand && bit(sandm, cid(q[2])) && ((s >> 20) & 3u) == 0) q[2] |= MOVED;
            if (bit(airm, cid(q[1]))) {
                const Cell side = q[1];
                q[1] = q[3];
                q[0] = side;
            } else {
                q[0] = q[3];
            }
            q[3] = a2;
            return true;
        }
        if (cr) {
            Cell b2 = q[1];
            const bool sand = bit(sandm, mb);
            if (sand && ub) b2 |= MOVED;
            if (sand && ((s >> 11) & 511u) == 0) b2 &= ~MOVED;
            if (sand && bit(sandm, cid(q[3])) && ((s >> 20) & 3u) == 0) q[3] |= MOVED;
            if (bit(airm, cid(q[0]))) {
                const Cell side = q[0];
                q[0] = q[2];
                q[1] = side;
            } else {
                q[1] = q[2];
            }
            q[2] = b2;
            return true;
        }
        bool gl = bit(gasm, cid(q[2])) && bit(airm, cid(q[1]));
        bool gr = bit(gasm, cid(q[3])) && bit(airm, cid(q[0]));
        if (gl && gr) {
            if ((s >> 22) & 1u) gl = false;
            else gr = false;
        }
        if (gl) {
            const Cell t = q[2];
            q[2] = q[1];
            q[1] = t;
            return true;
        }
        if (gr) {
            const Cell t = q[3];
            q[3] = q[0];
            q[0] = t;
            return true;
        }
    }

    // F: an active sand cell with neither its straight nor its in-block
    // diagonal move open stops (MOVED cleared)
#pragma unroll
    for (int col = 0; col < 2; col++) {
        const int o = 1 - col;
        const uint32_t mt = cid(q[col]);
        const int sr = nib(sinkt, mt);
        if (bit(sandm, mt) && sr > 0 && nib(pass, cid(q[col + 2])) >= sr && nib(pass, cid(q[o + 2])) >= sr)
            q[col] &= ~MOVED;
    }

    // H: per row, top then bottom.  Soup flows into a lighter non-solid cell
    // when it faces it (MOVED set: heading right), or turns with p=1/16
    // (r[24:27] top, r[28:31] bottom) and takes the new heading; otherwise
    // gas spreads into air with p=1/2
#pragma unroll
    for (int row = 0; row < 2; row++) {
        const int i = 2 * row, j = i + 1;
        const uint32_t ml = cid(q[i]), mr = cid(q[j]);
        const bool fl = nib(pass, mr) < nib(flowt, ml), fr = nib(pass, ml) < nib(flowt, mr);
        bool go;
        if (fl || fr) {
            const bool facing = fl ? (q[i] & MOVED) != 0 : (q[j] & MOVED) == 0;
            go = facing || ((r >> (24 + 4 * row)) & 15u) == 0;
            if (go) {
                if (fl) q[i] |= MOVED;
                else q[j] &= ~MOVED;
            }
        } else {
            go = ((bit(gasm, ml) && bit(airm, mr)) || (bit(gasm, mr) && bit(airm, ml))) && ((s >> (23 + row)) & 1u);
        }
        if (go) {
            const Cell t = q[i];
            q[i] = q[j];
            q[j] = t;
        }
    }
    return q[0] != o0 || q[1] != o1 || q[2] != o2 || q[3] != o3;
}

// ---------------------------------------------------------------------------
// Table-driven form of block_rule for the hot loop.
//
// Every decision of block_rule is a relation between two cells of the block
// (or a property of one cell).  pair_code(t, u) packs those relations for an
// ordered pair of material ids at iteration `it` into 9 bits; a workgroup
// builds the 3 x 256-entry table in LDS once, and a block then needs seven
// table reads instead of dozens of nibble lookups.  The outcome is expressed
// as an XOR mask on the identity permutation of the block (2 bits per
// destination) plus per-cell MOVED set/clear masks, applied once, so the
// divergent stages only combine lane masks.  block_rule (above) stays the
// readable statement of the same rule and handles reacting blocks.
// ---------------------------------------------------------------------------
constexpr uint32_t K_SINK = 1u << 0;     // t sinks into u
constexpr uint32_t K_GRISE = 1u << 1;    // u is active gas and t is air
constexpr uint32_t K_HSWAP = 1u << 2;    // the horizontal pair (t, u) may swap
constexpr uint32_t K_REACT = 1u << 3;    // t water above u lava, reactions on
constexpr uint32_t K_TSAND = 1u << 4;
constexpr uint32_t K_USAND = 1u << 5;
constexpr uint32_t K_TAIR = 1u << 6;
constexpr uint32_t K_UAIR = 1u << 7;
constexpr uint32_t K_TSANDACT = 1u << 8; // t is sand active in this iteration
constexpr uint32_t K_FLOWTU = 1u << 9;   // t, an active SOUP cell, flows sideways into u
constexpr uint32_t K_FLOWUT = 1u << 10;  // u flows sideways into t

__host__ __device__ __forceinline__ uint32_t pair_code(const MatTab &M, int it, uint32_t t, uint32_t u) {
    const int pt = nib(M.pass, t), pu = nib(M.pass, u);
    const int st = nib(M.sink[it], t);
    const int ft = nib(M.flow[it], t), fu = nib(M.flow[it], u);
    const bool gt = bit(M.gas[it], t), gu = bit(M.gas[it], u);
    const bool at = bit(M.air, t), au = bit(M.air, u);
    uint32_t c = 0;
    if (pu < st) c |= K_SINK;
    if (gu && at) c |= K_GRISE;
    if (pu < ft || pt < fu || (gt && au) || (gu && at)) c |= K_HSWAP;
    if (M.reactions && t == M.water && u == M.lava) c |= K_REACT;
    if (bit(M.sand, t)) c |= K_TSAND;
    if (bit(M.sand, u)) c |= K_USAND;
    if (at) c |= K_TAIR;
    if (au) c |= K_UAIR;
    if (bit(M.sand, t) && st > 0) c |= K_TSANDACT;
    if (pu < ft) c |= K_FLOWTU;
    if (pt < fu) c |= K_FLOWUT;
    return c;
}

// Decision of a non-reacting block, as one word:
//   bits 0-7   scatter mask: cell k moves to position k ^ ((y >> 2k) & 3)
//   bits 8-11  MOVED set on original cell k (bit 8 + k)
//   bits 12-15 MOVED cleared on original cell k (bit 12 + k)
// All intermediate values are 0/1 integers, so the stages combine with plain
// bit operations instead of divergent branches.  Same result as block_rule
// for blocks where no reaction fires.  GAS_H = false compiles the form for
// blocks with no K_GRISE bit in VL, VR, DL, DR and no K_HSWAP bit in HT, HB
// (the gas and horizontal terms are then zero).
template <bool GAS_H = true>
__host__ __device__ __forceinline__ uint32_t block_decide(uint32_t VL, uint32_t VR, uint32_t DL, uint32_t DR,
                                                          uint32_t HT, uint32_t HTR, uint32_t HB,
                                                          uint32_t movA, uint32_t movB, uint32_t movC,
                                                          uint32_t movD, uint32_t r, uint32_t s) {
    if (!GAS_H) {
        VL &= ~K_GRISE;
        VR &= ~K_GRISE;
        DL &= ~K_GRISE;
        DR &= ~K_GRISE;
        HT = 0;
        HB = 0;
    }
    const uint32_t sL = VL & 1u, sR = VR & 1u, dL = DL & 1u, dR = DR & 1u, htr = HTR & 1u;
    const uint32_t gL = (VL >> 1) & 1u, gR = (VR >> 1) & 1u, gdl = (DL >> 1) & 1u, gdr = (DR >> 1) & 1u;
    const uint32_t sandA = (VL >> 4) & 1u, sandB = (VR >> 4) & 1u;
    const uint32_t sandC = (VL >> 5) & 1u, sandD = (VR >> 5) & 1u;
    const uint32_t airA = (VL >> 6) & 1u, airB = (DR >> 6) & 1u, airC = (VL >> 7) & 1u;
    const uint32_t actA = (VL >> 8) & 1u, actB = (VR >> 8) & 1u;
    const uint32_t hT = (HT >> 2) & 1u, hB = (HB >> 2) & 1u;
    const uint32_t T1 = (r & 0xFu) == 0, T2 = (r & 0xF0u) == 0, T5 = (r & 0xC00u) == 0, T6 = (r & 0x3000u) == 0;
    const uint32_t T3 = (r >> 8) & 1u, T4 = (r >> 9) & 1u, T7 = (r >> 14) & 1u, T8 = (r >> 15) & 1u;
    const uint32_t UA = (r & 0xF0000u) == 0, UB = (r & 0xF00000u) == 0;  // stopped sand unsticks
    const uint32_t S1 = (s & 0x1Fu) == 0, S2 = (s & 0x3E0u) == 0, S4 = (s & 0xFF800u) == 0;
    const uint32_t S5 = (s & 0x300000u) == 0;
    const uint32_t S3 = (s >> 10) & 1u, S6 = (s >> 22) & 1u, S7 = (s >> 23) & 1u, S8 = (s >> 24) & 1u;

    // V: left column, then the right column seeing the left column's result
    const uint32_t vLs = sL & ((sandA & dL & T1) ^ 1u);
    const uint32_t vLg = (sL ^ 1u) & gL & ((airB & T7) ^ 1u);
    const uint32_t vL = vLs | vLg;
    const uint32_t sinksBnewC = vL ? htr : dR;           // after a left swap, new BL is old TL
    const uint32_t vRs = sR & ((sandB & sinksBnewC & T2) ^ 1u);
    const uint32_t airNewA = vL ? airC : airA;           // after a left swap, new TL is old BL
    const uint32_t vRg = (sR ^ 1u) & gR & ((airNewA & T8) ^ 1u);
    const uint32_t vR = vRs | vRg;
    const uint32_t vm = vL | vR;
    uint32_t set = (vLs & sandA & T3) | ((vRs & sandB & T4) << 1) | ((vLs & sandA & T5 & sandD) << 3);
    const uint32_t nudR = vRs & sandB & T6 & (vL ? sandA : sandC);
    set |= vL ? nudR : (nudR << 2);
    uint32_t y = (vL * 0x22u) | (vR * 0x88u);

    // D: sinking diagonals with sand friction, then gas diagonals
    const uint32_t nvm = vm ^ 1u;
    const uint32_t cl0 = nvm & dL, cr0 = nvm & dR;
    const uint32_t mvA = movA | UA, mvB = movB | UB;
    const uint32_t failA = cl0 & sandA & mvA & S1, failB = cr0 & sandB & mvB & S2;
    const uint32_t cl1 = cl0 & ((sandA ^ 1u) | (mvA & (failA ^ 1u)));
    const uint32_t cr1 = cr0 & ((sandB ^ 1u) | (mvB & (failB ^ 1u)));
    const uint32_t both = cl1 & cr1;
    const uint32_t cl = cl1 & ((both & S3) ^ 1u);
    const uint32_t cr = cr1 & ((both & (S3 ^ 1u)) ^ 1u);
    uint32_t clr = failA | (failB << 1) | (cl & sandA & S4) | ((cr & sandB & S4) << 1);
    set |= ((cl & sandA & sandC & S5) << 2) | ((cr & sandB & sandD & S5) << 3);
    set |= (cl & sandA & UA) | ((cr & sandB & UB) << 1);  // an unstuck slider moves (S4 may clear it)
    y |= cl * (0xC3u ^ (airB * 0x44u));  // swap TL/BR, or 3-cycle TL->BR->TR->TL when TR is air
    y |= cr * (0x3Cu ^ (airA * 0x11u));  // swap TR/BL, or 3-cycle TR->BL->TL->TR when TL is air
    const uint32_t none = nvm & (cl ^ 1u) & (cr ^ 1u);
    const uint32_t gl0 = none & gdr, gr0 = none & gdl;
    const uint32_t gb = gl0 & gr0;
    const uint32_t gl = gl0 & ((gb & S6) ^ 1u), gr = gr0 & ((gb & (S6 ^ 1u)) ^ 1u);
    y |= (gl * 0x3Cu) | (gr * 0xC3u);
    const uint32_t rest = none & (gl ^ 1u) & (gr ^ 1u);

    // F: an active sand top cell with no straight or in-block diagonal move stops
    clr |= (rest & actA & (sL ^ 1u) & (dL ^ 1u)) | ((rest & actB & (sR ^ 1u) & (dR ^ 1u)) << 1);
    // H: per row, a SOUP cell that can flow sideways moves if it faces the
    // other cell (MOVED set: heading right), else turns and moves with
    // p=1/16, taking the new heading; otherwise gas spreads with p=1/2
    const uint32_t fAB = (HT >> 9) & 1u, fBA = (HT >> 10) & 1u, fCD = (HB >> 9) & 1u, fDC = (HB >> 10) & 1u;
    const uint32_t lT = fAB | fBA, lB = fCD | fDC;
    const uint32_t TT = (r & 0xF000000u) == 0, TB = (r & 0xF0000000u) == 0;
    const uint32_t goT = rest & ((lT & ((fAB & movA) | (fBA & (movB ^ 1u)) | TT)) | ((lT ^ 1u) & hT & S7));
    const uint32_t goB = rest & ((lB & ((fCD & movC) | (fDC & (movD ^ 1u)) | TB)) | ((lB ^ 1u) & hB & S8));
    y ^= goT * 0x05u;
    y ^= goB * 0x50u;
    set |= (goT & fAB) | ((goB & fCD) << 2);
    clr |= ((goT & fBA) << 1) | ((goB & fDC) << 3);
    return y | (set << 8) | (clr << 12);
}

// True if no rule can change the block for any random draw: no pair
// relation, no active sand top cell whose MOVED flag could be cleared.
__host__ __device__ __forceinline__ bool block_inert(uint32_t VL, uint32_t VR, uint32_t DL, uint32_t DR,
                                                     uint32_t HT, uint32_t HB, uint32_t movA, uint32_t movB) {
    const uint32_t rel = (VL | VR | DL | DR | HT | HB) & (K_SINK | K_GRISE | K_HSWAP | K_REACT);
    return rel == 0 && ((((VL >> 8) & movA) | ((VR >> 8) & movB)) & 1u) == 0;
}

// Host form of the decision's effect on q (the kernel scatters to LDS).
template <typename Cell>
__host__ __device__ __forceinline__ void apply_decision(uint32_t dec, Cell q[4]) {
    Cell v[4];
    for (int k = 0; k < 4; k++) {
        v[k] = q[k];
        if ((dec >> (8 + k)) & 1u) v[k] |= MOVED;
        if ((dec >> (12 + k)) & 1u) v[k] &= ~MOVED;
    }
    for (int k = 0; k < 4; k++) q[k ^ ((dec >> (2 * k)) & 3u)] = v[k];
}

// Pair tests of the dirty-rect predicate (iteration gate taken at it = 0).
__device__ __forceinline__ bool pair_vert(const MatTab &M, uint64_t t, uint64_t u) {
    const uint32_t mt = cid(t), mu = cid(u);
    return nib(M.pass, mu) < nib(M.sink[0], mt) || (bit(M.gas[0], mu) && bit(M.air, mt)) ||
           (M.reactions && cbyte(t) == M.water && cbyte(u) == M.lava);
}
__device__ __forceinline__ bool pair_diag(const MatTab &M, uint64_t t, uint64_t u) {
    const uint32_t mt = cid(t), mu = cid(u);
    return nib(M.pass, mu) < nib(M.sink[0], mt) || (bit(M.gas[0], mu) && bit(M.air, mt));
}
__device__ __forceinline__ bool pair_horz(const MatTab &M, uint64_t l, uint64_t r) {
    const uint32_t ml = cid(l), mr = cid(r);
    return nib(M.pass, mr) < nib(M.flow[0], ml) || nib(M.pass, ml) < nib(M.flow[0], mr) ||
           (bit(M.gas[0], ml) && bit(M.air, mr)) || (bit(M.gas[0], mr) && bit(M.air, ml));
}
__device__ __forceinline__ bool single_unstable(const MatTab &M, uint64_t c) {
    // a SOUP cell's heading (its MOVED bit) changes only with a move
    return ((c & MOVED) && bit(M.sand, cid(c))) || (M.reactions && cbyte(c) == M.wood && ctemp(c) > M.burn);
}

} // namespace fsg
