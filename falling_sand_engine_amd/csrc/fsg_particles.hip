// fsg_particles.hip -- free-fall particles (CellData, game_utils/cells.h:15-36)
// and world::tickCells (world.cpp:2030-2189) on the device.
//
// The reference moves its particles one after another, and a particle that
// lands changes the cells the particles after it move through and search:
// particle i's outcome is a function of the grid and of the landings of the
// particles before it.  The device finds those outcomes as a fixed point.
// Every particle moves in parallel against the grid plus a view of the
// landings the previous round found, in which a landing counts for particle
// i only if an earlier particle made it (a hash of landing cell -> lowest
// index).  A round whose landings equal the previous round's has the
// outcomes the sequential order gives: by induction over the list, each
// particle saw exactly the landings before it.  After any round the particles
// before the first one whose landing changed are already final, so the next
// round moves only the rest; a round with no landing change ends the pass.
// A list whose rounds keep changing (a long chain of landings that depend on
// each other) hands its unfinished part to the batch walk below, which keeps
// the order by construction.
//
// Batch walk: one workgroup walks the list in batches of 256: every lane
// moves its particle against the grid as the earlier batches left it, and
// records the bounding box of every cell it read (its path, the cell it
// started in, the spiral it searched).  A lane whose box holds the landing
// cell of an earlier lane of the batch might have seen something else had it
// run after that lane; the first such lane ends the batch, the lanes before
// it commit, and the next batch starts at it.
#include "fsg_internal.hpp"

#include <algorithm>
#include <cmath>

namespace {

constexpr int PT = 256;  // particles per batch (threads of the one workgroup)
constexpr int NSPIRAL = 1024;

// tickCells' landing spiral (world.cpp:2113-2139), walked in place: from
// (0, 0) heading (0, -1), turning left where x == y, or x < 0 and x == -y,
// or x > 0 and x == 1 - y.  Its 1024 steps stay within [-15, 16]^2.  (Walking
// it costs a few VALU a step; a table read per step is a memory latency in
// the divergent search loop.)
struct Spiral {
    int x = 0, y = 0, dx = 0, dy = -1;
    __device__ __forceinline__ void next() {
        if (x == y || (x < 0 && x == -y) || (x > 0 && x == 1 - y)) {
            const int t = dx;
            dx = -dy;
            dy = t;
        }
        x += dx;
        y += dy;
    }
};

struct PState {
    fsg_particle p;
    int kind;         // 0 keep, 1 drop, 2 land
    int land;         // cell index of a landing
    int bx0, by0, bx1, by1;  // bounding box of the cells read
};

__device__ __forceinline__ void grow_box(PState &s, int x, int y) {
    s.bx0 = min(s.bx0, x);
    s.by0 = min(s.by0, y);
    s.bx1 = max(s.bx1, x);
    s.by1 = max(s.by1, y);
}

__device__ __forceinline__ bool is_air(const uint64_t *cells, uint32_t airmask, int i) {
    return (airmask >> (uint32_t)(cells[i] & 15u)) & 1u;
}

// Landings seen by particle `self`: keys[h] = landing cell (EMPTY: free slot),
// vals[h] = lowest index of a particle that landed there with a non-air tile.
// keys == nullptr: no landings.
constexpr uint32_t EMPTY = 0xFFFFFFFFu;
struct View {
    const uint32_t *keys, *vals;
    uint32_t mask, self;
};

__device__ __forceinline__ uint32_t slot_of(uint32_t cell, uint32_t mask) { return fsg::lowbias32(cell) & mask; }

// An air cell of the grid that no earlier particle has filled.
__device__ __forceinline__ bool free_cell(const uint64_t *cells, uint32_t airmask, const View &v, int i) {
    if (!is_air(cells, airmask, i)) return false;
    if (!v.keys) return true;
    for (uint32_t h = slot_of((uint32_t)i, v.mask);; h = (h + 1) & v.mask) {  // the table is at most half full
        const uint32_t k = v.keys[h];
        if (k == EMPTY) return true;
        if (k == (uint32_t)i) return v.vals[h] >= v.self;
    }
}

// Which 8 x 8 blocks of the grid hold an air cell, found on demand: two bits
// per block in one word (bit 2k: known, bit 2k + 1: air), so one atomicOr
// publishes both.  Cleared at the start of a pass; the grid only loses air
// cells during a pass, so a block once known to hold none never gains one.
struct AirMap {
    uint32_t *bits;  // null: no map
    int wpr;         // words per block row (16 blocks a word)
};

__device__ bool block_has_air(const AirMap &am, const uint64_t *cells, int W, int H, uint32_t airmask, int bx,
                              int by) {
    uint32_t *wd = am.bits + (size_t)by * am.wpr + (bx >> 4);
    const uint32_t sh = 2u * (uint32_t)(bx & 15);
    const uint32_t v = *wd;
    if ((v >> sh) & 1u) return (v >> (sh + 1)) & 1u;
    bool air = false;
    const int x0 = bx * 8, y0 = by * 8, x1 = min(x0 + 8, W), y1 = min(y0 + 8, H);
    if (x1 - x0 == 8 && y1 - y0 == 8 && (W & 1) == 0) {
        // a whole block of an even-width grid: four rounds of eight 16-byte
        // loads issued together (two rows each)
#pragma unroll 1
        for (int r2 = 0; r2 < 4; r2++) {
            uint4 q[8];
#pragma unroll
            for (int k = 0; k < 8; k++)
                q[k] = *reinterpret_cast<const uint4 *>(cells + (size_t)(y0 + 2 * r2 + (k >> 2)) * W + x0 + 2 * (k & 3));
#pragma unroll
            for (int k = 0; k < 8; k++) air |= ((airmask >> (q[k].x & 15u)) | (airmask >> (q[k].z & 15u))) & 1u;
        }
    } else {
        for (int y = y0; y < y1; y++)
            for (int x = x0; x < x1; x++) air |= (airmask >> (uint32_t)(cells[(size_t)y * W + x] & 15u)) & 1u;
    }
    atomicOr(wd, (1u | (air ? 2u : 0u)) << sh);
    return air;
}

// One particle's tick against the grid (the reference's loop body).
__device__ void move_particle(PState &s, const uint64_t *cells, int W, int H, uint32_t airmask, int4 zone,
                              const View &view, const AirMap &am) {
#pragma clang fp contract(off)
    fsg_particle c = s.p;
    s.kind = 0;
    s.land = -1;
    s.bx0 = s.by0 = 1 << 30;
    s.bx1 = s.by1 = -(1 << 30);
    if ((c.flags & FSG_PARTICLE_TEMPORARY) && c.lifetime <= 0) {
        s.kind = 1;
        return;
    }
    if (c.target_force != 0.0f) {
        const float tdx = c.target_x - c.x, tdy = c.target_y - c.y;
        const float nf = sqrtf(tdx * tdx + tdy * tdy);
        c.vx += tdx / nf * c.target_force;
        c.vy += tdy / nf * c.target_force;
        if (nf < 100.0f) {
            c.vx *= 0.95f;
            c.vy *= 0.95f;
        }
    }
    const int lx = (int)c.x, ly = (int)c.y;
    if (c.x < 0.0f || (int)c.x >= W || c.y < 0.0f || (int)c.y >= H) {
        s.kind = 1;
        return;
    }
    if (lx < zone.x || ly < zone.y || lx >= zone.z || ly >= zone.w) {  // outside the tick zone: kept as it is
        s.p = c;
        return;
    }
    c.vx += c.ax;
    c.vy += c.ay;
    const int div = (int)((fabsf(c.vx) + fabsf(c.vy)) + 1.0f);
    const float dvx = c.vx / (float)div, dvy = c.vy / (float)div;
    bool done = false;
    for (int k = 0; k < div && !done; k++) {
        c.x += dvx;
        c.y += dvy;
        if (c.x < 0.0f || (int)c.x >= W || c.y < 0.0f || (int)c.y >= H) {
            s.kind = 1;
            return;
        }
        const int px = (int)c.x, py = (int)c.y;
        if (c.flags & FSG_PARTICLE_PHASE) continue;
        grow_box(s, px, py);
        if (free_cell(cells, airmask, view, px + py * W)) continue;
        if (c.in_object_state <= 1) c.in_object_state = 2;  // no OBJECT cells on this grid
        done = true;
        if (c.flags & FSG_PARTICLE_TEMPORARY) {
            s.kind = 1;
            return;
        }
        grow_box(s, lx, ly);
        if (!free_cell(cells, airmask, view, lx + ly * W)) {
            // SB offsets at a time: their cells are loaded together, then
            // tested in spiral order.  After the first SB, a 5 x 5 mask of the
            // 8 x 8 blocks the spiral can reach says which hold an air cell;
            // offsets in the others are not loaded (no cell there is free),
            // and with no air block the search ends at once.
            constexpr int SB = 16;
            uint32_t m25 = 0;
            int bx0 = 0, by0 = 0;
            bool masked = false;
            Spiral sp;
            for (int k0 = 0; k0 < NSPIRAL; k0 += SB) {
                if (k0 == SB && am.bits) {
                    // the spiral reads cells of [(int)(x - 15), (int)(x + 16)] x
                    // [(int)(y - 15), (int)(y + 16)] (the float sums are monotone in
                    // the offset), at most 5 blocks each way
                    const int xa = max((int)(c.x + -15.0f), 0), xb = min((int)(c.x + 16.0f), W - 1);
                    const int ya = max((int)(c.y + -15.0f), 0), yb = min((int)(c.y + 16.0f), H - 1);
                    bx0 = xa >> 3;
                    by0 = ya >> 3;
                    for (int by = by0; by <= (yb >> 3); by++)
                        for (int bx = bx0; bx <= (xb >> 3); bx++)
                            if (block_has_air(am, cells, W, H, airmask, bx, by)) m25 |= 1u << ((by - by0) * 5 + (bx - bx0));
                    grow_box(s, xa, ya);  // the rest of the search reads within the rectangle
                    grow_box(s, xb, yb);
                    if (!m25) break;
                    masked = true;
                }
                int at[SB];
                uint32_t m[SB];
#pragma unroll
                for (int u = 0; u < SB; u++) {
                    const int cx = (int)(c.x + (float)sp.x);
                    const int cy = (int)(c.y + (float)sp.y);
                    sp.next();
                    bool in = cx >= 0 && cy >= 0 && cx < W && cy < H;  // the reference reads out of range here
                    if (masked && in) in = (m25 >> (((cy >> 3) - by0) * 5 + ((cx >> 3) - bx0))) & 1u;
                    at[u] = in ? cx + cy * W : -1;
                    m[u] = in ? (uint32_t)cells[cx + cy * W] : 0u;
                    if (in && !masked) grow_box(s, cx, cy);
                }
#pragma unroll
                for (int u = 0; u < SB; u++) {
                    if (at[u] < 0 || !((airmask >> (m[u] & 15u)) & 1u)) continue;
                    if (free_cell(cells, airmask, view, at[u])) {
                        s.kind = 2;
                        s.land = at[u];
                        return;
                    }
                }
            }
            c.vy = -4.0f;
            c.y -= 16.0f;
        } else {
            s.kind = 2;
            s.land = lx + ly * W;
            return;
        }
    }
    if (!done && c.lifetime > 0) c.lifetime--;
    s.p = c;
}

struct PArgs {
    fsg_particle *p;       // the list, updated in place
    uint8_t *alive;        // per particle: kept after the pass
    uint32_t start;        // first particle not yet final
    uint32_t n;
    uint64_t *cells;
    int W, H, tiles_x;
    uint32_t airmask;
    int4 zone;             // tick zone [x, z) x [y, w)
    uint32_t *flags;       // chunk flags the next tick reads (null: no skipping)
    unsigned long long *landed;
    uint32_t *airbits;     // the pass's air-block map
    int air_wpr;
};

__global__ __launch_bounds__(PT) void k_particles(PArgs a) {
    __shared__ int land_at[PT];
    __shared__ int first_bad, next;
    const int t = threadIdx.x;
    uint32_t base = a.start;
    unsigned long long nl = 0;
    const View none{nullptr, nullptr, 0u, 0u};
    const AirMap am{a.airbits, a.air_wpr};
    while (base < a.n) {  // base is the same in every thread
        const uint32_t i = base + (uint32_t)t;
        PState s;
        const bool mine = i < a.n;
        if (mine) {
            s.p = a.p[i];
            move_particle(s, a.cells, a.W, a.H, a.airmask, a.zone, none, am);
        }
        land_at[t] = (mine && s.kind == 2) ? s.land : -1;
        if (t == 0) first_bad = PT;
        __syncthreads();
        // the first lane whose reads cover an earlier lane's landing
        bool bad = false;
        if (mine && s.bx0 <= s.bx1) {
            for (int j = 0; j < t && !bad; j++) {
                const int c = land_at[j];
                if (c < 0) continue;
                const int cy = c / a.W, cx = c - cy * a.W;
                bad = cx >= s.bx0 && cx <= s.bx1 && cy >= s.by0 && cy <= s.by1;
            }
        }
        if (bad) atomicMin(&first_bad, t);
        __syncthreads();
        const int fb = first_bad;
        if (mine && t < fb) {
            if (s.kind == 2) {
                const fsg_cell &tl = s.p.tile;
                a.cells[s.land] = (uint64_t)tl.mat | ((uint64_t)(uint16_t)tl.temp << 16) | ((uint64_t)tl.rgba << 32);
                if (a.flags) {
                    const int cy = s.land / a.W, cx = s.land - cy * a.W;
                    atomicOr(&a.flags[(cy / fsg::TH) * a.tiles_x + cx / fsg::TW], fsg::F_WAKE_ALL);
                }
                nl++;
            }
            a.alive[i] = s.kind == 0 && !(s.p.y > (float)a.H);  // erase_if(y > height), world.cpp:2190
            if (s.kind == 0) a.p[i] = s.p;
        }
        if (t == 0) next = (int)min((uint32_t)fb, a.n - base);
        __syncthreads();
        base += (uint32_t)next;
        __threadfence_block();  // the landings are seen by the next batch's loads
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) nl += __shfl_xor(nl, off, 64);
    if ((t & 63) == 0 && nl) atomicAdd(a.landed, nl);
}

// ---- fixed-point rounds ------------------------------------------------------
struct PRes {
    fsg_particle p;  // state if kept
    int32_t kind;    // 0 keep, 1 drop, 2 land
    int32_t land;    // landing cell
    int32_t bx0, by0, bx1, by1;  // bounding box of the cells it read
};

__global__ void k_u32_fill(uint32_t *p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_ph_clear(uint32_t *keys, uint32_t *vals, uint32_t cap, uint32_t val0, uint32_t *cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) {
        keys[i] = EMPTY;
        vals[i] = val0;
    }
    if (i < 4 && cnt) cnt[i] = i == 1 ? 0xFFFFFFFFu : 0u;  // [0] changes, [1] first change, [2] changed cells
}

// Adds the landings of particles [0, n): key[i] is a cell or -1.  MAX keeps
// the highest index per cell (the last writer), else the lowest.
template <bool MAX>
__global__ void k_ph_insert(const int32_t *key, uint32_t n, uint32_t *keys, uint32_t *vals, uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || key[i] < 0) return;
    const uint32_t c = (uint32_t)key[i];
    for (uint32_t h = slot_of(c, mask);; h = (h + 1) & mask) {  // capacity >= 2n: a slot is always found
        const uint32_t prev = atomicCAS(&keys[h], EMPTY, c);
        if (prev == EMPTY || prev == c) {
            if (MAX) atomicMax(&vals[h], i);
            else atomicMin(&vals[h], i);
            return;
        }
    }
}

__device__ __forceinline__ uint32_t ph_lookup(const uint32_t *keys, const uint32_t *vals, uint32_t mask, uint32_t c) {
    for (uint32_t h = slot_of(c, mask);; h = (h + 1) & mask) {
        const uint32_t k = keys[h];
        if (k == c) return vals[h];
        if (k == EMPTY) return EMPTY;
    }
}

struct RoundArgs {
    const fsg_particle *p;  // the list as the pass found it
    PRes *res;
    int32_t *vkey;          // per particle: its non-air landing cell of the last round, or -1
    int32_t *wkey;          // per particle: its landing cell (any tile), or -1
    uint32_t start, n;
    const uint64_t *cells;
    int W, H;
    uint32_t airmask;
    int4 zone;
    View view;              // self is set per particle
    uint32_t *cnt;          // [0] landings that changed, [1] first particle whose landing changed,
                            // [2] cells appended to chg
    const int32_t *prev_chg;  // cells whose landing changed in the last round (null: move every particle)
    uint32_t prev_nchg;
    int32_t *chg;           // this round's changed cells
    AirMap am;
};

// One round: particles [start, n) against the grid and the view.  A
// particle none of whose read cells changed its landing in the last round
// reads what it read then, so its outcome stands and it is not moved again.
constexpr uint32_t MAX_CHG = 4096;  // changed cells a round tests against (more: every particle moves)

__global__ __launch_bounds__(256) void k_ph_round(RoundArgs a) {
    __shared__ int2 cl[MAX_CHG];  // the changed cells as (x, y), staged once per workgroup
    const uint32_t i = a.start + blockIdx.x * blockDim.x + threadIdx.x;
    if (a.prev_chg) {
        for (uint32_t k = threadIdx.x; k < a.prev_nchg; k += blockDim.x) {
            const int c = a.prev_chg[k];
            const int cy = c / a.W;
            cl[k] = make_int2(c - cy * a.W, cy);
        }
        __syncthreads();
        if (i >= a.n) return;
        const PRes &o = a.res[i];
        bool hit = false;
        if (o.bx0 <= o.bx1) {
            for (uint32_t k = 0; k < a.prev_nchg; k++) {
                const int2 q = cl[k];
                hit |= q.x >= o.bx0 && q.x <= o.bx1 && q.y >= o.by0 && q.y <= o.by1;
            }
        }
        if (!hit) return;
    }
    if (i >= a.n) return;
    PState s;
    s.p = a.p[i];
    View v = a.view;
    v.self = i;
    move_particle(s, a.cells, a.W, a.H, a.airmask, a.zone, v, a.am);
    PRes r;
    r.p = s.p;
    r.kind = s.kind;
    r.land = s.land;
    r.bx0 = s.bx0;
    r.by0 = s.by0;
    r.bx1 = s.bx1;
    r.by1 = s.by1;
    a.res[i] = r;
    const bool solid = s.kind == 2 && !((a.airmask >> (uint32_t)(s.p.tile.mat & 15u)) & 1u);
    const int32_t k = solid ? s.land : -1;
    const int32_t old = a.vkey[i];
    if (k != old) {
        atomicAdd(&a.cnt[0], 1u);
        atomicMin(&a.cnt[1], i);
        a.vkey[i] = k;
        const uint32_t at = atomicAdd(&a.cnt[2], (old >= 0 ? 1u : 0u) + (k >= 0 ? 1u : 0u));
        uint32_t o2 = at;
        if (old >= 0) a.chg[o2++] = old;
        if (k >= 0) a.chg[o2] = k;
    }
    a.wkey[i] = s.kind == 2 ? s.land : -1;
}

struct ApplyArgs {
    fsg_particle *p;
    const PRes *res;
    uint8_t *alive;
    uint32_t n;             // particles [0, n) are final
    uint64_t *cells;
    int W, H, tiles_x;
    uint32_t *flags;
    const uint32_t *keys, *vals;  // last writer per landing cell
    uint32_t mask;
    unsigned long long *landed;
};

// Final outcomes: the last particle landing in a cell writes it (an earlier
// one can only have carried an air tile, which a later landing replaces).
__global__ __launch_bounds__(256) void k_ph_apply(ApplyArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long nl = 0;
    if (i < a.n) {
        const PRes r = a.res[i];
        if (r.kind == 2) {
            nl = 1;
            if (ph_lookup(a.keys, a.vals, a.mask, (uint32_t)r.land) == i) {
                const fsg_cell &tl = r.p.tile;
                a.cells[r.land] = (uint64_t)tl.mat | ((uint64_t)(uint16_t)tl.temp << 16) | ((uint64_t)tl.rgba << 32);
                if (a.flags) {
                    const int cy = r.land / a.W, cx = r.land - cy * a.W;
                    atomicOr(&a.flags[(cy / fsg::TH) * a.tiles_x + cx / fsg::TW], fsg::F_WAKE_ALL);
                }
            }
        }
        a.alive[i] = r.kind == 0 && !(r.p.y > (float)a.H);  // erase_if(y > height), world.cpp:2190
        if (r.kind == 0) a.p[i] = r.p;
    }
    for (int off = 32; off > 0; off >>= 1) nl += __shfl_xor(nl, off, 64);
    if ((threadIdx.x & 63) == 0 && nl) atomicAdd(a.landed, nl);
}

// ---- survivors in list order ---------------------------------------------------
// Counts per 1024-particle block, an exclusive scan of the counts in one
// workgroup, then each block writes its survivors from its offset.
__global__ __launch_bounds__(1024) void k_pc_count(const uint8_t *alive, uint32_t n, uint32_t *blk) {
    __shared__ uint32_t wsum[16];
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    const uint64_t m = __ballot(i < n && alive[i]);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < 16; k++) t += wsum[k];
        blk[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void k_pc_scan(uint32_t *blk, uint32_t nb, uint32_t *total) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? blk[i] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
            const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nb) blk[i] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void k_pc_scatter(const fsg_particle *in, const uint8_t *alive, uint32_t n,
                                                     const uint32_t *blk, fsg_particle *out) {
    __shared__ uint32_t wsum[16];
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool keep = i < n && alive[i];
    const uint64_t m = __ballot(keep);
    if (lane == 0) wsum[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = blk[blockIdx.x];
    for (int k = 0; k < wv; k++) off += wsum[k];
    if (keep) out[off + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = in[i];
}

using fsg_impl::dev_alloc;
using fsg_impl::dev_free;

// Grows the list buffers to hold `want` particles, keeping the list.
int grow_lists(fsg_world *w, size_t want) {
    if (w->part_cap >= want) return FSG_OK;
    fsg_ctx *c = w->ctx;
    const size_t nc = std::max(want, w->part_cap * 2);
    void *a = nullptr, *b = nullptr, *f = nullptr;
    if (dev_alloc(c, &a, nc * sizeof(fsg_particle))) return FSG_ERR_NOMEM;
    if (dev_alloc(c, &b, nc * sizeof(fsg_particle))) {
        dev_free(a);
        return FSG_ERR_NOMEM;
    }
    if (dev_alloc(c, &f, nc)) {
        dev_free(a);
        dev_free(b);
        return FSG_ERR_NOMEM;
    }
    if (w->npart) HIPCHK(hipMemcpyAsync(a, w->d_part, w->npart * sizeof(fsg_particle), hipMemcpyDeviceToDevice, c->stream));
    dev_free(w->d_part);
    dev_free(w->d_part2);
    dev_free(w->d_part_alive);
    w->d_part = static_cast<fsg_particle *>(a);
    w->d_part2 = static_cast<fsg_particle *>(b);
    w->d_part_alive = static_cast<uint8_t *>(f);
    w->part_cap = nc;
    return FSG_OK;
}

// Round buffers for `n` particles, carved from one block: results, the two
// landing keys, the compaction's block offsets, counters, and the landing
// hash (capacity the power of two >= 2n, so it is at most half full).
struct RoundBufs {
    PRes *res;
    int32_t *vkey, *wkey, *chg[2];
    uint32_t *blk, *cnt, *keys, *vals, *air;
    uint32_t hcap, air_words;
    int air_wpr;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

int round_bufs(fsg_world *w, size_t n, RoundBufs &b) {
    uint32_t hcap = 1024;
    while (hcap < 2 * n) hcap <<= 1;
    const size_t nb = n / 1024 + 2;
    const int nbx = (w->d.W + 7) / 8, nby = (w->d.H + 7) / 8;
    b.air_wpr = (nbx + 15) / 16;
    b.air_words = (uint32_t)((size_t)b.air_wpr * nby);
    const size_t sz = align256(n * sizeof(PRes)) + 2 * align256(n * 4) + 2 * align256(n * 8) + align256(nb * 4) + 256 +
                      2 * align256((size_t)hcap * 4) + align256((size_t)b.air_words * 4);
    if (w->pround_cap < sz) {
        void *q = nullptr;
        if (dev_alloc(w->ctx, &q, sz)) return FSG_ERR_NOMEM;
        dev_free(w->d_pround);
        w->d_pround = q;
        w->pround_cap = sz;
    }
    char *c = static_cast<char *>(w->d_pround);
    b.res = reinterpret_cast<PRes *>(c);
    c += align256(n * sizeof(PRes));
    b.vkey = reinterpret_cast<int32_t *>(c);
    c += align256(n * 4);
    b.wkey = reinterpret_cast<int32_t *>(c);
    c += align256(n * 4);
    for (int k = 0; k < 2; k++) {  // each particle adds at most two cells per round
        b.chg[k] = reinterpret_cast<int32_t *>(c);
        c += align256(n * 8);
    }
    b.blk = reinterpret_cast<uint32_t *>(c);
    c += align256(nb * 4);
    b.cnt = reinterpret_cast<uint32_t *>(c);
    c += 256;
    b.keys = reinterpret_cast<uint32_t *>(c);
    c += align256((size_t)hcap * 4);
    b.vals = reinterpret_cast<uint32_t *>(c);
    c += align256((size_t)hcap * 4);
    b.air = reinterpret_cast<uint32_t *>(c);
    b.hcap = hcap;
    return FSG_OK;
}

// Rounds before a list whose landings keep changing goes to the batch walk.
constexpr int MAX_ROUNDS = 24;

}  // namespace

extern "C" int fsg_world_particles_add(fsg_world *w, const fsg_particle *p, size_t n) {
    if (!w || (n && !p)) return FSG_ERR_INVALID;
    for (size_t i = 0; i < n; i++) {
        const uint16_t m = p[i].tile.mat;
        if ((m & 0xFE00u) || (uint32_t)(m & 0xFFu) >= w->cfg.n_materials) return FSG_ERR_INVALID;
        if (!std::isfinite(p[i].x) || !std::isfinite(p[i].y) || !std::isfinite(p[i].vx) || !std::isfinite(p[i].vy) ||
            !std::isfinite(p[i].ax) || !std::isfinite(p[i].ay) || !std::isfinite(p[i].target_x) ||
            !std::isfinite(p[i].target_y) || !std::isfinite(p[i].target_force) || p[i].in_object_state > 2 ||
            (p[i].flags & ~(FSG_PARTICLE_PHASE | FSG_PARTICLE_TEMPORARY)))
            return FSG_ERR_INVALID;
    }
    if (n == 0) return FSG_OK;
    if (w->npart + n > 0xFFFFFFFFull) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    if (int rc = grow_lists(w, w->npart + n)) return rc;
    HIPCHK(hipMemcpyAsync(w->d_part + w->npart, p, n * sizeof(fsg_particle), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    w->npart += n;
    return FSG_OK;
}

extern "C" int fsg_world_particles_tick(fsg_world *w, size_t *alive, size_t *landed) {
    if (!w) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    const uint32_t n = (uint32_t)w->npart;
    unsigned long long nl = 0;
    if (n) {
        if (!w->d_part_cnt) {
            void *q = nullptr;
            if (dev_alloc(c, &q, 16)) return FSG_ERR_NOMEM;
            w->d_part_cnt = static_cast<unsigned long long *>(q);
        }
        RoundBufs b;
        if (int rc = round_bufs(w, n, b)) return rc;
        hipStream_t st = c->stream;
        uint32_t airmask = 0;
        for (uint32_t m = 0; m < w->cfg.n_materials; m++)
            if (w->cfg.materials[m].type == FSG_AIR) airmask |= 1u << m;
        const int4 zone = make_int4(w->d.zx0, w->d.zy0, w->d.zx1, w->d.zy1);
        uint64_t *cells = w->d.buf[w->d.tick & 1];
        // landings wake their chunks for the next tick (no flags: it processes every chunk)
        uint32_t *flags = w->flags_valid ? w->d.flags[(w->d.tick + 1) & 1] : nullptr;
        HIPCHK(hipMemsetAsync(w->d_part_cnt, 0, 16, st));
        HIPCHK(hipMemsetAsync(b.vkey, 0xFF, (size_t)n * 4, st));  // -1: no landing yet
        const unsigned hg = (b.hcap + 255) / 256;
        k_u32_fill<<<(b.air_words + 255) / 256, 256, 0, st>>>(b.air, b.air_words, 0u);
        RoundArgs ra{w->d_part, b.res, b.vkey, b.wkey, 0u, n, cells, w->d.W, w->d.H, airmask, zone,
                     View{nullptr, nullptr, b.hcap - 1, 0u}, b.cnt, nullptr, 0u, b.chg[0], AirMap{b.air, b.air_wpr}};
        uint32_t final_n = 0;  // particles [0, final_n) have their outcomes
        for (int round = 0; round < MAX_ROUNDS; round++) {
            k_ph_clear<<<hg, 256, 0, st>>>(b.keys, b.vals, b.hcap, EMPTY, b.cnt);
            if (round > 0) {
                k_ph_insert<false><<<(n + 255) / 256, 256, 0, st>>>(b.vkey, n, b.keys, b.vals, b.hcap - 1);
                ra.view.keys = b.keys;
                ra.view.vals = b.vals;
            }
            ra.start = final_n;
            ra.chg = b.chg[round & 1];
            k_ph_round<<<(n - final_n + 255) / 256, 256, 0, st>>>(ra);
            HIPCHK(hipGetLastError());
            uint32_t h[3];
            HIPCHK(hipMemcpyAsync(h, b.cnt, 12, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (h[0] == 0) {  // no landing changed: every outcome is the sequential one
                final_n = n;
                break;
            }
            final_n = h[1];  // the particles before the first changed landing are final
            // the next round moves only particles that read a changed cell
            // (a long list of changes costs more to test than to move all)
            ra.prev_chg = h[2] <= MAX_CHG ? b.chg[round & 1] : nullptr;
            ra.prev_nchg = h[2];
        }
        if (final_n > 0) {
            // the last particle landing in a cell writes it
            k_ph_clear<<<hg, 256, 0, st>>>(b.keys, b.vals, b.hcap, 0u, nullptr);
            k_ph_insert<true><<<(final_n + 255) / 256, 256, 0, st>>>(b.wkey, final_n, b.keys, b.vals, b.hcap - 1);
            ApplyArgs aa{w->d_part, b.res, w->d_part_alive, final_n, cells, w->d.W, w->d.H, w->d.tiles_x, flags,
                         b.keys, b.vals, b.hcap - 1, w->d_part_cnt};
            k_ph_apply<<<(final_n + 255) / 256, 256, 0, st>>>(aa);
        }
        if (final_n < n) {
            // the rest, in order, by the batch walk over the grid the prefix left
            PArgs a;
            a.p = w->d_part;
            a.alive = w->d_part_alive;
            a.start = final_n;
            a.n = n;
            a.cells = cells;
            a.W = w->d.W;
            a.H = w->d.H;
            a.tiles_x = w->d.tiles_x;
            a.zone = zone;
            a.airmask = airmask;
            a.flags = flags;
            a.landed = w->d_part_cnt;
            a.airbits = b.air;
            a.air_wpr = b.air_wpr;
            k_particles<<<1, PT, 0, st>>>(a);
        }
        const uint32_t nb = (n + 1023) / 1024;
        k_pc_count<<<nb, 1024, 0, st>>>(w->d_part_alive, n, b.blk);
        k_pc_scan<<<1, 1024, 0, st>>>(b.blk, nb, reinterpret_cast<uint32_t *>(w->d_part_cnt + 1));
        k_pc_scatter<<<nb, 1024, 0, st>>>(w->d_part, w->d_part_alive, n, b.blk, w->d_part2);
        HIPCHK(hipGetLastError());
        unsigned long long h[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(h, w->d_part_cnt, 16, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        nl = h[0];
        std::swap(w->d_part, w->d_part2);  // the survivors, in order, are the list
        w->npart = (size_t)(uint32_t)h[1];
    }
    if (alive) *alive = w->npart;
    if (landed) *landed = (size_t)nl;
    return FSG_OK;
}

extern "C" int fsg_world_particles_get(fsg_world *w, fsg_particle *out, size_t cap, size_t *n) {
    if (!w || (cap && !out)) return FSG_ERR_INVALID;
    fsg_ctx *c = w->ctx;
    HIPCHK(hipSetDevice(c->device));
    const size_t k = std::min(cap, w->npart);
    if (k) {
        HIPCHK(hipMemcpyAsync(out, w->d_part, k * sizeof(fsg_particle), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (n) *n = w->npart;
    return FSG_OK;
}
