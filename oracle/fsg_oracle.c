/*
 * fsg_oracle.c -- deterministic CPU restatement of the falling-sand tick.
 *
 * TEST INFRASTRUCTURE ONLY (see fsg_oracle.h): the checker for the HIP path,
 * never linked into the product.
 *
 * The reference tick (ME::world::tick, source/engine/world.cpp:1036-1948) is
 * a sequential bottom-to-top scan per 128x128 chunk with global rand() on 16
 * threads, so its trajectories are not reproducible.  This file defines the
 * deterministic tick that replaces it and that the GPU matches bit for bit:
 *
 *   tick t = cell_iter phases (GlobalDEF::cell_iter, global.lua:47), phase
 *   `it` being a Margolus pass over disjoint 2x2 blocks whose offset comes
 *   from the global phase index g = t*cell_iter + it; then, if the cadence
 *   says so, one integer temperature step (game.cpp:2157-2159).
 *
 * Each block is updated by orc_block_rule from its own four cells and two
 * 32-bit random words keyed on (seed, g, block position).  Blocks never
 * overlap inside a phase, so the result is independent of block order.
 * The rules restate the reference's per-cell SAND/SOUP/GAS rules within the
 * block; every rule cites the lines it follows.
 */
#include "fsg_oracle.h"
#include <stdlib.h>
#include <string.h>

#define GOLDEN 0x9E3779B97F4A7C15ull
#define WALL_ID 15
#define MOVED ((uint64_t)FSG_MOVED)

/* ----------------------------------------------------------------------- */
/* RNG                                                                      */
/* ----------------------------------------------------------------------- */

/* splitmix64 (S. Vigna, public-domain reference implementation): the output
 * of next() when the state before the call is `state`. */
uint64_t orc_splitmix64(uint64_t state) {
    uint64_t z = state + GOLDEN;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* lowbias32 (C. Wellons, hash-prospector): a 32-bit bijective mixer. */
uint32_t orc_lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* Its published inverse. */
uint32_t orc_lowbias32_inv(uint32_t x) {
    x ^= x >> 16;
    x *= 0x43021123u;
    x ^= x >> 15 ^ x >> 30;
    x *= 0x1d69e2a5u;
    x ^= x >> 16;
    return x;
}

/* Key of global phase g for a world seeded with `seed`.  Replaces the global
 * rand() of world.cpp (passim) and RNG_Create's time(NULL) seed
 * (game_utils/rng.cpp:8-19). */
uint64_t orc_phase_key(uint64_t seed, uint64_t g) {
    return orc_splitmix64(orc_splitmix64(seed) ^ g);
}

/* Random words of the block whose top-left cell is (x0, y0); x0, y0 >= -1. */
void orc_block_hash(uint64_t key, int x0, int y0, uint32_t *r, uint32_t *s) {
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
    uint32_t pos = (uint32_t)(x0 + 1) | ((uint32_t)(y0 + 1) << 16);
    uint32_t a = orc_lowbias32(pos ^ k0);
    *r = a;
    *s = orc_lowbias32(a ^ k1);
}

/* ----------------------------------------------------------------------- */
/* Material properties (Material, game_datastruct.hpp:130-168)              */
/* ----------------------------------------------------------------------- */
typedef struct props {
    uint8_t type[16], rank[16], iters[16];
    int16_t add_temp[16];
    uint32_t rgba[16];
} props;

static void build_props(const fsg_config *cfg, props *P) {
    for (int i = 0; i < 16; i++) {
        /* ids without a table entry behave like the wall: inert solid */
        P->type[i] = FSG_SOLID;
        P->rank[i] = 15;
        P->iters[i] = 0;
        P->add_temp[i] = 0;
        P->rgba[i] = 0;
    }
    for (uint32_t i = 0; i < cfg->n_materials && i < FSG_MAX_MATERIALS; i++) {
        P->type[i] = cfg->materials[i].type;
        P->rank[i] = cfg->materials[i].rank;
        P->iters[i] = cfg->materials[i].iterations;
        P->add_temp[i] = cfg->materials[i].add_temp;
        P->rgba[i] = cfg->materials[i].rgba;
    }
}

static inline int mid(uint64_t c) { return (int)(c & 0xFF); }
static inline int temp_of(uint64_t c) { return (int)(int16_t)(uint16_t)(c >> 16); }

/* Rank a cell offers to a sinking neighbour: SOLID never yields
 * (canMoveBelow's "below != SOLID", world.cpp:1206). */
static inline int pass_rank(const props *P, int m) {
    return P->type[m] == FSG_SOLID ? 15 : P->rank[m];
}
/* Rank with which an active SAND/SOUP cell sinks; 0 = cannot sink.  The
 * iteration gate is world.cpp:1093-1096. */
static inline int sink_rank(const props *P, int m, int it) {
    int t = P->type[m];
    return ((t == FSG_SAND || t == FSG_SOUP) && it < P->iters[m]) ? P->rank[m] : 0;
}
/* Rank with which an active SOUP cell flows sideways; 0 = does not. */
static inline int flow_rank(const props *P, int m, int it) {
    return (P->type[m] == FSG_SOUP && it < P->iters[m]) ? P->rank[m] : 0;
}
static inline int gas_act(const props *P, int m, int it) {
    return P->type[m] == FSG_GAS && it < P->iters[m];
}
static inline int is_air(const props *P, int m) { return P->type[m] == FSG_AIR; }
static inline int is_sand(const props *P, int m) { return P->type[m] == FSG_SAND; }

/* New cell of material m (TilesCreate*, game_datastruct.cpp:304-574):
 * keeps the temperature (world.cpp:1186-1187), takes the material's base
 * colour, clears the flags. */
static inline uint64_t transform(const props *P, uint64_t c, int m) {
    return (uint64_t)(uint32_t)m | (c & 0xFFFF0000ull) | ((uint64_t)P->rgba[m] << 32);
}

#define SWAP(a, b) do { uint64_t _t = (a); (a) = (b); (b) = _t; } while (0)

/* ----------------------------------------------------------------------- */
/* The block rule                                                           */
/* ----------------------------------------------------------------------- */
/* q[0]=TL (x0,y0)  q[1]=TR (x0+1,y0)  q[2]=BL (x0,y0+1)  q[3]=BR (x0+1,y0+1)
 * Random bits: r[0:3]/r[4:7] sand vertical skip L/R, r[8]/r[9] MOVED on fall
 * L/R, r[10:11]/r[12:13] nudge L/R, r[14]/r[15] gas rise skip L/R,
 * r[16:19]/r[20:23] stopped sand unsticks TL/TR;
 * r[24:27]/r[28:31] a SOUP cell facing away turns, top/bottom row;
 * s[0:4]/s[5:9] slide refused L/R, s[10] slide side choice, s[11:19] MOVED
 * cleared after a slide, s[20:21] slide nudge, s[22] gas diagonal choice,
 * s[23]/s[24] gas spreads in the top/bottom row.
 * MOVED is sand's sliding flag and a SOUP cell's heading. */
static int block_rule(const props *P, const fsg_config *cfg, int it, uint32_t r, uint32_t s,
                      uint64_t q[4]) {
    const uint64_t o0 = q[0], o1 = q[1], o2 = q[2], o3 = q[3];
#define CHANGED() (q[0] != o0 || q[1] != o1 || q[2] != o2 || q[3] != o3)

    /* --- R: reactions ------------------------------------------------- */
    if (cfg->reactions) {
        int fired = 0;
        const int W = cfg->mat_water, L = cfg->mat_lava;
        /* Water directly above lava: the water becomes steam and every lava
         * cell of the block becomes stone (world.cpp:1519-1537: water ->
         * steam, lava below and lava in x-1..x+1, y..y+2 -> obsidian; stone
         * plays obsidian's role and the neighbourhood is the block). */
        int wl = mid(q[0]) == W && mid(q[2]) == L;
        int wr = mid(q[1]) == W && mid(q[3]) == L;
        if (wl || wr) {
            if (wl) q[0] = transform(P, q[0], cfg->mat_steam);
            if (wr) q[1] = transform(P, q[1], cfg->mat_steam);
            for (int i = 0; i < 4; i++)
                if (mid(q[i]) == L) q[i] = transform(P, q[i], cfg->mat_stone);
            fired = 1;
        }
        /* Wood above the burn temperature becomes steam and trades places
         * with the first air cell of the block, leaving air where the wood
         * was (REACT_TEMPERATURE_ABOVE form, world.cpp:1193-1200). */
        for (int i = 0; i < 4; i++) {
            if (mid(q[i]) == cfg->mat_wood && temp_of(q[i]) > cfg->wood_burn_temp) {
                q[i] = transform(P, q[i], cfg->mat_steam);
                for (int j = 0; j < 4; j++)
                    if (is_air(P, mid(q[j]))) { SWAP(q[i], q[j]); break; }
                fired = 1;
            }
        }
        if (fired) return CHANGED(); /* reacted cells do not also move */
    }

    /* --- V: vertical moves, left column then right -------------------- */
    int vm = 0;
    for (int col = 0; col < 2; col++) {
        const int o = 1 - col;
        uint64_t t = q[col], u = q[col + 2];
        const int mt = mid(t), mu = mid(u);
        const int sr = sink_rank(P, mt, it);
        if (pass_rank(P, mu) < sr) {
            /* SAND/SOUP falls or sinks through a lighter non-solid cell
             * (canMoveBelow, world.cpp:1206; swap, 1227-1239). */
            const int sand = is_sand(P, mt);
            /* Sand with a free diagonal skips the straight fall 1/16 of the
             * time (rand()%20 in world.cpp:1217). */
            if (sand && pass_rank(P, mid(q[o + 2])) < sr && ((r >> (4 * col)) & 15u) == 0)
                continue;
            /* The faller is marked moving with p=1/2 (world.cpp:1231-1233). */
            if (sand && ((r >> (8 + col)) & 1u)) t |= MOVED;
            q[col] = u;
            q[col + 2] = t;
            vm = 1;
            /* A SAND cell beside the landing spot is nudged into motion with
             * p=1/4 (world.cpp:1242-1266). */
            if (sand && ((r >> (10 + 2 * col)) & 3u) == 0 && is_sand(P, mid(q[o + 2])))
                q[o + 2] |= MOVED;
        } else if (gas_act(P, mu, it) && is_air(P, mt)) {
            /* GAS rises into air, skipping with p=1/2 when the other upper
             * cell is air too (world.cpp:1569-1585). */
            if (is_air(P, mid(q[o])) && ((r >> (14 + col)) & 1u)) continue;
            q[col] = u;
            q[col + 2] = t;
            vm = 1;
        }
    }
    if (vm) return 1;

    /* --- D: diagonal moves ------------------------------------------- */
    {
        const int ma = mid(q[0]), mb = mid(q[1]);
        int cl = pass_rank(P, mid(q[3])) < sink_rank(P, ma, it); /* TL -> BR */
        int cr = pass_rank(P, mid(q[2])) < sink_rank(P, mb, it); /* TR -> BL */
        /* Sand friction (world.cpp:1612-1700): a sand cell slides while
         * MOVED is set, and refuses with p=1/32 (shouldMove, rand()%(2*slip)
         * at 1656), which also stops it (moved = false, 1695-1700).  A
         * stopped sand cell whose diagonal is open comes unstuck with p=1/16
         * (r[16:19] == 0 for TL, r[20:23] == 0 for TR) and slides with MOVED
         * set: the block-local form of the friction pillar check, which
         * unsticks a stopped cell above a drop (world.cpp:1616-1644) and
         * cannot be restated exactly because it reads 10 rows below the
         * cell. */
        const int ua = ((r >> 16) & 15u) == 0, ub = ((r >> 20) & 15u) == 0;
        if (cl && is_sand(P, ma)) {
            if (!(q[0] & MOVED) && !ua) cl = 0;
            else if ((s & 31u) == 0) { cl = 0; q[0] &= ~MOVED; }
        }
        if (cr && is_sand(P, mb)) {
            if (!(q[1] & MOVED) && !ub) cr = 0;
            else if (((s >> 5) & 31u) == 0) { cr = 0; q[1] &= ~MOVED; }
        }
        if (cl && cr) { if ((s >> 10) & 1u) cl = 0; else cr = 0; }
        if (cl) {
            uint64_t a2 = q[0];
            const int sand = is_sand(P, ma);
            if (sand && ua) a2 |= MOVED; /* unstuck (world.cpp:1636) */
            /* moved cleared with p=1/512 after a slide (1/(20*slip), 1688). */
            if (sand && ((s >> 11) & 511u) == 0) a2 &= ~MOVED;
            /* the cell below the slider is nudged with p=1/4 (1658-1671). */
            if (sand && is_sand(P, mid(q[2])) && ((s >> 20) & 3u) == 0) q[2] |= MOVED;
            /* Side cell air: the displaced diagonal cell moves up beside and
             * the slider's old place becomes air; else a plain swap
             * (world.cpp:1676-1686). */
            if (is_air(P, mid(q[1]))) {
                uint64_t side = q[1];
                q[1] = q[3];
                q[3] = a2;
                q[0] = side;
            } else {
                q[0] = q[3];
                q[3] = a2;
            }
            return 1;
        }
        if (cr) {
            uint64_t b2 = q[1];
            const int sand = is_sand(P, mb);
            if (sand && ub) b2 |= MOVED;
            if (sand && ((s >> 11) & 511u) == 0) b2 &= ~MOVED;
            if (sand && is_sand(P, mid(q[3])) && ((s >> 20) & 3u) == 0) q[3] |= MOVED;
            if (is_air(P, mid(q[0]))) {
                uint64_t side = q[0];
                q[0] = q[2];
                q[2] = b2;
                q[1] = side;
            } else {
                q[1] = q[2];
                q[2] = b2;
            }
            return 1;
        }
        /* GAS rises diagonally into air (world.cpp:1799-1819). */
        int gl = gas_act(P, mid(q[2]), it) && is_air(P, mid(q[1])); /* BL -> TR */
        int gr = gas_act(P, mid(q[3]), it) && is_air(P, mid(q[0])); /* BR -> TL */
        if (gl && gr) { if ((s >> 22) & 1u) gl = 0; else gr = 0; }
        if (gl) { SWAP(q[2], q[1]); return 1; }
        if (gr) { SWAP(q[3], q[0]); return 1; }
    }

    /* --- F: a sand cell with no move left in the block stops -----------
     * (stoppedByFriction || !(canMoveBelowL || canMoveBelowR) -> moved =
     * false, world.cpp:1646-1652). */
    for (int col = 0; col < 2; col++) {
        const int o = 1 - col;
        const int mt = mid(q[col]);
        const int sr = sink_rank(P, mt, it);
        if (is_sand(P, mt) && sr > 0 && !(pass_rank(P, mid(q[col + 2])) < sr) &&
            !(pass_rank(P, mid(q[o + 2])) < sr))
            q[col] &= ~MOVED;
    }

    /* --- H: horizontal moves, top row then bottom row -----------------
     * SOUP flows sideways into a lighter non-solid cell (the discrete SOUP
     * rule, world.cpp:1839-1861) and GAS spreads into air (1862-1882).
     * A SOUP cell keeps a heading in its MOVED bit (set: right, clear:
     * left).  Facing the cell it can flow into, it moves; facing away, it
     * turns and moves with p=1/16 (r[24:27] top row, r[28:31] bottom row).
     * The reference's scan moves a liquid front every pass, toward the one
     * side that is free (world.cpp:1839-1861 in bottom-up, left-to-right
     * order); a Margolus pair seen twice in a row would undo a plain swap,
     * and a fair coin per pair makes the front a random walk, so the heading
     * is what carries a front forward.  GAS spreads with p=1/2
     * (s[23]/s[24]). */
    for (int row = 0; row < 2; row++) {
        const int i = 2 * row, j = i + 1;
        const int ml = mid(q[i]), mr = mid(q[j]);
        const int fl = pass_rank(P, mr) < flow_rank(P, ml, it); /* left flows right */
        const int fr = pass_rank(P, ml) < flow_rank(P, mr, it); /* right flows left */
        if (fl || fr) {
            const int facing = fl ? (q[i] & MOVED) != 0 : (q[j] & MOVED) == 0;
            const int turn = ((r >> (24 + 4 * row)) & 15u) == 0;
            if (facing || turn) {
                if (fl) q[i] |= MOVED;
                else q[j] &= ~MOVED;
                SWAP(q[i], q[j]);
            }
        } else {
            const int g = (gas_act(P, ml, it) && is_air(P, mr)) || (gas_act(P, mr, it) && is_air(P, ml));
            if (g && ((s >> (23 + row)) & 1u)) SWAP(q[i], q[j]);
        }
    }
    return CHANGED();
#undef CHANGED
}

int orc_block_rule(const fsg_config *cfg, int it, uint32_t r, uint32_t s, uint64_t q[4]) {
    props P;
    build_props(cfg, &P);
    return block_rule(&P, cfg, it, r, s, q);
}

/* ----------------------------------------------------------------------- */
/* Tick                                                                     */
/* ----------------------------------------------------------------------- */
/* Block offset of global phase g: (0,0), (1,1), (1,0), (0,1) repeating, so
 * the row offset alternates every phase. */
static inline void phase_offset(uint64_t g, int *ox, int *oy) {
    static const int OX[4] = {0, 1, 1, 0}, OY[4] = {0, 1, 0, 1};
    *ox = OX[g & 3];
    *oy = OY[g & 3];
}

static const uint64_t WALL_CELL = (uint64_t)WALL_ID;

static inline uint64_t cell_at(const uint64_t *c, int W, int H, int x, int y) {
    if (x < 0 || y < 0 || x >= W || y >= H) return WALL_CELL; /* getTile OOB, world.cpp:1000 */
    return c[(size_t)y * W + x];
}

/* Tick zone (World::tickZone, world.hpp:138, set in game.cpp:1629-1636):
 * world::tick visits only the chunks whose origin lies in the zone
 * (world.cpp:1073-1074), and a cell it visits may move into cells outside.
 * Here a block acts iff at least one of its four cells lies in the zone
 * zone[0..3] = x, y, w, h, so cells leave the zone across any edge, cells
 * of the one-cell ring outside it move only together with a zone cell, and
 * a zone covering the world is the same as no zone (NULL). */
static int block_in_zone(const int32_t *zone, int x0, int y0) {
    return !zone || (x0 + 1 >= zone[0] && x0 < zone[0] + zone[2] && y0 + 1 >= zone[1] && y0 < zone[1] + zone[3]);
}
static int cell_in_zone(const int32_t *zone, int x, int y) {
    return !zone || (x >= zone[0] && x < zone[0] + zone[2] && y >= zone[1] && y < zone[1] + zone[3]);
}

static void run_phase(uint64_t *cells, int W, int H, const props *P, const fsg_config *cfg,
                      uint64_t seed, uint64_t tick, int it, const int32_t *zone) {
    const uint64_t g = tick * (uint64_t)cfg->cell_iter + (uint64_t)it;
    const uint64_t key = orc_phase_key(seed, g);
    int ox, oy;
    phase_offset(g, &ox, &oy);
    for (int y0 = oy ? -1 : 0; y0 < H; y0 += 2) {
        for (int x0 = ox ? -1 : 0; x0 < W; x0 += 2) {
            uint64_t q[4] = {cell_at(cells, W, H, x0, y0), cell_at(cells, W, H, x0 + 1, y0),
                             cell_at(cells, W, H, x0, y0 + 1),
                             cell_at(cells, W, H, x0 + 1, y0 + 1)};
            if (!block_in_zone(zone, x0, y0)) continue;
            uint32_t r, s;
            orc_block_hash(key, x0, y0, &r, &s);
            if (!block_rule(P, cfg, it, r, s, q)) continue;
            /* setTile semantics: out-of-world writes are dropped
             * (world.cpp:1004-1008); walls never change anyway. */
            const int xs[4] = {x0, x0 + 1, x0, x0 + 1}, ys[4] = {y0, y0, y0 + 1, y0 + 1};
            for (int k = 0; k < 4; k++)
                if (xs[k] >= 0 && ys[k] >= 0 && xs[k] < W && ys[k] < H)
                    cells[(size_t)ys[k] * W + xs[k]] = q[k];
        }
    }
}

static void temperature_step(uint64_t *cells, int W, int H, const props *P, const int32_t *zone) {
    /* Integer 5-point diffusion with coefficient 1/8, floor rounding
     * (arithmetic shift); an out-of-world neighbour counts as the cell
     * itself.  Replaces world::tickTemperature's float 9-point mean
     * (world.cpp:1950-2004); add_temp is Material::addTemp (1988).  Only
     * zone cells change (the loops of world.cpp:1953-1954); their stencil
     * reads neighbours outside the zone as they are. */
    int16_t *nt = (int16_t *)malloc((size_t)W * H * sizeof(int16_t));
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            const uint64_t c = cells[(size_t)y * W + x];
            const int T = temp_of(c);
            if (!cell_in_zone(zone, x, y)) {
                nt[(size_t)y * W + x] = (int16_t)T;
                continue;
            }
            const int tu = y > 0 ? temp_of(cells[(size_t)(y - 1) * W + x]) : T;
            const int td = y < H - 1 ? temp_of(cells[(size_t)(y + 1) * W + x]) : T;
            const int tl = x > 0 ? temp_of(cells[(size_t)y * W + x - 1]) : T;
            const int tr = x < W - 1 ? temp_of(cells[(size_t)y * W + x + 1]) : T;
            int v = T + ((tu + td + tl + tr - 4 * T) >> 3) + P->add_temp[mid(c)];
            if (v > 32767) v = 32767;
            if (v < -32768) v = -32768;
            nt[(size_t)y * W + x] = (int16_t)v;
        }
    }
    for (size_t i = 0; i < (size_t)W * H; i++)
        cells[i] = (cells[i] & 0xFFFFFFFF0000FFFFull) | ((uint64_t)(uint16_t)nt[i] << 16);
    free(nt);
}

static int temp_runs(const fsg_config *cfg, uint64_t tick) {
    return cfg->temp_every > 0 && (tick % cfg->temp_every) == cfg->temp_phase;
}

static int zone_ok(const int32_t *zone, int W, int H) {
    return !zone || (zone[0] >= 0 && zone[1] >= 0 && zone[2] >= 1 && zone[3] >= 1 && zone[2] <= W - zone[0] &&
                     zone[3] <= H - zone[1]);
}

int orc_temperature_zone(uint64_t *cells, int W, int H, const fsg_config *cfg, const int32_t *zone) {
    if (!cells || W <= 0 || H <= 0 || !zone_ok(zone, W, H)) return FSG_ERR_INVALID;
    props P;
    build_props(cfg, &P);
    temperature_step(cells, W, H, &P, zone);
    return 0;
}

int orc_temperature(uint64_t *cells, int W, int H, const fsg_config *cfg) {
    return orc_temperature_zone(cells, W, H, cfg, NULL);
}

int orc_tick_zone(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed, uint64_t tick,
                  const int32_t *zone) {
    if (!cells || W <= 0 || H <= 0 || cfg->cell_iter == 0 || !zone_ok(zone, W, H)) return FSG_ERR_INVALID;
    props P;
    build_props(cfg, &P);
    for (int it = 0; it < (int)cfg->cell_iter; it++) run_phase(cells, W, H, &P, cfg, seed, tick, it, zone);
    if (temp_runs(cfg, tick)) temperature_step(cells, W, H, &P, zone);
    return 0;
}

int orc_tick(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed, uint64_t tick) {
    return orc_tick_zone(cells, W, H, cfg, seed, tick, NULL);
}

int orc_run_zone(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed, uint64_t tick0,
                 int nticks, const int32_t *zone) {
    for (int t = 0; t < nticks; t++) {
        int rc = orc_tick_zone(cells, W, H, cfg, seed, tick0 + (uint64_t)t, zone);
        if (rc) return rc;
    }
    return 0;
}

int orc_run(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed, uint64_t tick0,
            int nticks) {
    return orc_run_zone(cells, W, H, cfg, seed, tick0, nticks, NULL);
}

/* ----------------------------------------------------------------------- */
/* Dirty-rect predicate                                                     */
/* ----------------------------------------------------------------------- */
/* A block can only change if it holds a pair of cells that some rule could
 * act on (vertical, diagonal or horizontal) or a single cell that can change
 * by itself (a sand cell's set MOVED flag, burning wood).  The iteration gate is taken
 * at it = 0, where every material with iterations > 0 is active. */
static int pair_v(const props *P, const fsg_config *cfg, uint64_t t, uint64_t u) {
    const int mt = mid(t), mu = mid(u);
    return pass_rank(P, mu) < sink_rank(P, mt, 0) || (gas_act(P, mu, 0) && is_air(P, mt)) ||
           (cfg->reactions && mt == cfg->mat_water && mu == cfg->mat_lava);
}
static int pair_d(const props *P, uint64_t t, uint64_t u) {
    const int mt = mid(t), mu = mid(u);
    return pass_rank(P, mu) < sink_rank(P, mt, 0) || (gas_act(P, mu, 0) && is_air(P, mt));
}
static int pair_h(const props *P, uint64_t l, uint64_t r) {
    const int ml = mid(l), mr = mid(r);
    return pass_rank(P, mr) < flow_rank(P, ml, 0) || pass_rank(P, ml) < flow_rank(P, mr, 0) ||
           (gas_act(P, ml, 0) && is_air(P, mr)) || (gas_act(P, mr, 0) && is_air(P, ml));
}

int orc_region_unstable(const uint64_t *cells, int W, int H, const fsg_config *cfg, int x0,
                        int y0, int x1, int y1, int cx0, int cy0, int cx1, int cy1) {
    props P;
    build_props(cfg, &P);
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > W) x1 = W;
    if (y1 > H) y1 = H;
    for (int y = y0; y < y1; y++) {
        for (int x = x0; x < x1; x++) {
            const uint64_t c = cells[(size_t)y * W + x];
            if ((c & MOVED) && is_sand(&P, mid(c))) return 1; /* a SOUP cell's heading alone changes nothing */
            if (cfg->reactions && mid(c) == cfg->mat_wood && temp_of(c) > cfg->wood_burn_temp)
                return 1;
            if (x + 1 < x1 && pair_h(&P, c, cells[(size_t)y * W + x + 1])) return 1;
            if (y + 1 < y1) {
                if (pair_v(&P, cfg, c, cells[(size_t)(y + 1) * W + x])) return 1;
                if (x + 1 < x1 && pair_d(&P, c, cells[(size_t)(y + 1) * W + x + 1])) return 1;
                if (x - 1 >= x0 && pair_d(&P, c, cells[(size_t)(y + 1) * W + x - 1])) return 1;
            }
        }
    }
    if (cfg->temp_every > 0) {
        if (cx0 < 0) cx0 = 0;
        if (cy0 < 0) cy0 = 0;
        if (cx1 > W) cx1 = W;
        if (cy1 > H) cy1 = H;
        for (int y = cy0; y < cy1; y++) {
            for (int x = cx0; x < cx1; x++) {
                const uint64_t c = cells[(size_t)y * W + x];
                const int T = temp_of(c);
                const int tu = y > 0 ? temp_of(cells[(size_t)(y - 1) * W + x]) : T;
                const int td = y < H - 1 ? temp_of(cells[(size_t)(y + 1) * W + x]) : T;
                const int tl = x > 0 ? temp_of(cells[(size_t)y * W + x - 1]) : T;
                const int tr = x < W - 1 ? temp_of(cells[(size_t)y * W + x + 1]) : T;
                if (((tu + td + tl + tr - 4 * T) >> 3) + P.add_temp[mid(c)] != 0) return 1;
            }
        }
    }
    return 0;
}

/* ----------------------------------------------------------------------- */
/* Rigid-body pixels                                                        */
/* ----------------------------------------------------------------------- */
/* Q16.16 to int, toward zero: the (int) cast of game.cpp:1774-1775. */
static int q16_trunc(int64_t v) { return (int)(v >= 0 ? v >> 16 : -((-v) >> 16)); }

/* Pixel (tx, ty) of a body lands on
 *   wx = (int)((tx*c - (ty+1)*s + x) / 2^16),  wy = (int)((tx*s + (ty+1)*c + y) / 2^16)
 * (quotients truncated toward zero) in Q16.16 integer arithmetic: the float
 * rotate of game.cpp:1774-1775 with exact products. */
int orc_body_pixel_xy(const fsg_body *b, int tx, int ty, int *wx, int *wy) {
    const int64_t X = b->pos_x_q16, Y = b->pos_y_q16, C = b->cos_q16, S = b->sin_q16;
    const int64_t fx = (int64_t)tx * C - (int64_t)(ty + 1) * S + X;
    const int64_t fy = (int64_t)tx * S + (int64_t)(ty + 1) * C + Y;
    *wx = q16_trunc(fx);
    *wy = q16_trunc(fy);
    return 0;
}

/* Target offsets tried in order (checkDirs, game.cpp:1766, 1907). */
static const int DIRX[5] = {0, 1, -1, 0, 0};
static const int DIRY[5] = {0, 0, 0, 1, -1};

static int body_of_pixel(const fsg_body *bodies, int nbodies, size_t p, int *tx, int *ty) {
    for (int b = 0; b < nbodies; b++) {
        const size_t n = (size_t)bodies[b].mat_w * (size_t)bodies[b].mat_h;
        if (p >= bodies[b].pixel_offset && p < bodies[b].pixel_offset + n) {
            const size_t k = p - bodies[b].pixel_offset;
            *tx = (int)(k % (size_t)bodies[b].mat_w);
            *ty = (int)(k / (size_t)bodies[b].mat_w);
            return b;
        }
    }
    return -1;
}

/* Stamp (game.cpp:1760-1815) in 5 claim rounds: in round d every unplaced
 * non-air pixel of an enabled body proposes its d-th offset; a target is
 * available if no pixel claimed it before and its cell is AIR, SAND or SOUP;
 * among proposers of one target the lowest pixel index wins.  A SAND/SOUP
 * occupant is displaced (the reference ejects it as a CellData particle) and
 * reported in pixel order. */
int orc_stamp(uint64_t *cells, int W, int H, const fsg_config *cfg, const fsg_body *bodies,
              int nbodies, const uint64_t *pixels, size_t npixels, uint8_t *placed,
              uint32_t *n_placed, uint32_t *n_displaced, fsg_cell_edit *displaced_out,
              uint32_t displaced_cap, uint32_t *hits_sand, uint32_t *hits_soup) {
    props P;
    build_props(cfg, &P);
    uint8_t *claimed = (uint8_t *)calloc((size_t)W * H, 1);
    uint64_t *disp_cell = (uint64_t *)calloc(npixels ? npixels : 1, sizeof(uint64_t));
    int32_t *disp_xy = (int32_t *)calloc(npixels ? 2 * npixels : 1, sizeof(int32_t));
    uint8_t *disp = (uint8_t *)calloc(npixels ? npixels : 1, 1);
    memset(placed, 0, npixels);
    if (hits_sand) memset(hits_sand, 0, sizeof(uint32_t) * (size_t)nbodies);
    if (hits_soup) memset(hits_soup, 0, sizeof(uint32_t) * (size_t)nbodies);
    uint32_t np = 0, nd = 0;
    for (int d = 0; d < 5; d++) {
        for (size_t p = 0; p < npixels; p++) {
            int tx, ty;
            const int b = body_of_pixel(bodies, nbodies, p, &tx, &ty);
            if (b < 0 || !bodies[b].enabled || placed[p]) continue;
            if (is_air(&P, mid(pixels[p]))) continue;
            int wx, wy;
            orc_body_pixel_xy(&bodies[b], tx, ty, &wx, &wy);
            wx += DIRX[d];
            wy += DIRY[d];
            if (wx < 0 || wy < 0 || wx >= W || wy >= H) continue;
            const size_t ci = (size_t)wy * W + wx;
            if (claimed[ci]) continue;
            const int tt = P.type[mid(cells[ci])];
            if (tt != FSG_AIR && tt != FSG_SAND && tt != FSG_SOUP) continue;
            claimed[ci] = 1;
            if (tt != FSG_AIR) {
                disp[p] = 1;
                disp_cell[p] = cells[ci];
                disp_xy[2 * p] = wx;
                disp_xy[2 * p + 1] = wy;
                if (tt == FSG_SAND && hits_sand) hits_sand[b]++;
                if (tt == FSG_SOUP && hits_soup) hits_soup[b]++;
            }
            cells[ci] = pixels[p];
            placed[p] = 1;
            np++;
        }
    }
    for (size_t p = 0; p < npixels; p++) {
        if (!disp[p]) continue;
        if (displaced_out && nd < displaced_cap) {
            displaced_out[nd].x = disp_xy[2 * p];
            displaced_out[nd].y = disp_xy[2 * p + 1];
            memcpy(&displaced_out[nd].cell, &disp_cell[p], 8);
        }
        nd++;
    }
    if (n_placed) *n_placed = np;
    if (n_displaced) *n_displaced = nd;
    free(claimed);
    free(disp_cell);
    free(disp_xy);
    free(disp);
    return 0;
}

/* Unstamp (game.cpp:1896-1983) in 5 claim rounds: a pixel takes back the
 * cell at its d-th offset if that cell has the pixel's material id and was
 * not claimed before; lowest pixel index wins.  The cell becomes air that
 * keeps its temperature.  A pixel that found nothing is lost (becomes air)
 * when the cell under it is air (game.cpp:1958-1964). */
int orc_unstamp(uint64_t *cells, int W, int H, const fsg_config *cfg, const fsg_body *bodies,
                int nbodies, uint64_t *pixels, size_t npixels) {
    props P;
    build_props(cfg, &P);
    const uint64_t air_id = cfg->mat_air;
    uint8_t *claimed = (uint8_t *)calloc((size_t)W * H, 1);
    uint8_t *found = (uint8_t *)calloc(npixels ? npixels : 1, 1);
    for (int d = 0; d < 5; d++) {
        for (size_t p = 0; p < npixels; p++) {
            int tx, ty;
            const int b = body_of_pixel(bodies, nbodies, p, &tx, &ty);
            if (b < 0 || !bodies[b].enabled || found[p]) continue;
            if (is_air(&P, mid(pixels[p]))) continue;
            int wx, wy;
            orc_body_pixel_xy(&bodies[b], tx, ty, &wx, &wy);
            wx += DIRX[d];
            wy += DIRY[d];
            if (wx < 0 || wy < 0 || wx >= W || wy >= H) continue;
            const size_t ci = (size_t)wy * W + wx;
            if (claimed[ci]) continue;
            if (mid(cells[ci]) != mid(pixels[p])) continue;
            claimed[ci] = 1;
            pixels[p] = cells[ci];
            cells[ci] = air_id | (cells[ci] & 0xFFFF0000ull);
            found[p] = 1;
        }
    }
    for (size_t p = 0; p < npixels; p++) {
        int tx, ty;
        const int b = body_of_pixel(bodies, nbodies, p, &tx, &ty);
        if (b < 0 || !bodies[b].enabled || found[p]) continue;
        if (is_air(&P, mid(pixels[p]))) continue;
        int wx, wy;
        orc_body_pixel_xy(&bodies[b], tx, ty, &wx, &wy);
        if (wx < 0 || wy < 0 || wx >= W || wy >= H) continue;
        if (is_air(&P, mid(cells[(size_t)wy * W + wx]))) pixels[p] = air_id;
    }
    free(claimed);
    free(found);
    return 0;
}

/* The reference's own stamp loop (game.cpp:1768-1812), as a pin for the
 * claim-round form above: bodies in order, and inside a body tx outer, ty
 * inner (pixel-major, column by column); each non-air pixel takes the first
 * checkDirs offset whose cell is in the world and AIR, SAND or SOUP, at
 * once, before the next pixel is tried.  Nothing marks a cell as taken, so a
 * later pixel may land on an earlier SAND or SOUP typed body pixel and eject
 * it.  placed[p] = 1 for a placed pixel; displaced cells are reported in
 * ejection order. */
int orc_refstamp(uint64_t *cells, int W, int H, const fsg_config *cfg, const fsg_body *bodies, int nbodies,
                 const uint64_t *pixels, size_t npixels, uint8_t *placed, uint32_t *n_placed,
                 uint32_t *n_displaced, fsg_cell_edit *displaced_out, uint32_t displaced_cap) {
    props P;
    build_props(cfg, &P);
    memset(placed, 0, npixels);
    uint32_t np = 0, nd = 0;
    for (int b = 0; b < nbodies; b++) {
        const fsg_body *B = &bodies[b];
        if (!B->enabled) continue;
        for (int tx = 0; tx < B->mat_w; tx++)
            for (int ty = 0; ty < B->mat_h; ty++) {
                const size_t p = B->pixel_offset + (size_t)ty * B->mat_w + tx;
                if (p >= npixels || is_air(&P, mid(pixels[p]))) continue;
                int wx, wy;
                orc_body_pixel_xy(B, tx, ty, &wx, &wy);
                for (int d = 0; d < 5; d++) {
                    const int x = wx + DIRX[d], y = wy + DIRY[d];
                    if (x < 0 || y < 0 || x >= W || y >= H) continue;
                    const size_t ci = (size_t)y * W + x;
                    const int tt = P.type[mid(cells[ci])];
                    if (tt != FSG_AIR && tt != FSG_SAND && tt != FSG_SOUP) continue;
                    if (tt != FSG_AIR) {
                        if (displaced_out && nd < displaced_cap) {
                            displaced_out[nd].x = x;
                            displaced_out[nd].y = y;
                            memcpy(&displaced_out[nd].cell, &cells[ci], 8);
                        }
                        nd++;
                    }
                    cells[ci] = pixels[p];
                    if (!placed[p]) np++;
                    placed[p] = 1;
                    break;
                }
            }
    }
    if (n_placed) *n_placed = np;
    if (n_displaced) *n_displaced = nd;
    return 0;
}

/* The reference's unstamp loop (game.cpp:1908-1965) in its own order: tx
 * outer, ty inner; a pixel takes the first checkDirs cell in the world with
 * its material id (MaterialInstance::id is the material id,
 * game_datastruct.cpp:304-309), which becomes air keeping its temperature
 * (the build's air, as orc_unstamp); a pixel that finds none is lost when
 * the cell at its own position is air. */
int orc_refunstamp(uint64_t *cells, int W, int H, const fsg_config *cfg, const fsg_body *bodies, int nbodies,
                   uint64_t *pixels, size_t npixels) {
    props P;
    build_props(cfg, &P);
    const uint64_t air_id = cfg->mat_air;
    for (int b = 0; b < nbodies; b++) {
        const fsg_body *B = &bodies[b];
        if (!B->enabled) continue;
        for (int tx = 0; tx < B->mat_w; tx++)
            for (int ty = 0; ty < B->mat_h; ty++) {
                const size_t p = B->pixel_offset + (size_t)ty * B->mat_w + tx;
                if (p >= npixels || is_air(&P, mid(pixels[p]))) continue;
                int wx, wy;
                orc_body_pixel_xy(B, tx, ty, &wx, &wy);
                int found = 0;
                for (int d = 0; d < 5 && !found; d++) {
                    const int x = wx + DIRX[d], y = wy + DIRY[d];
                    if (x < 0 || y < 0 || x >= W || y >= H) continue;
                    const size_t ci = (size_t)y * W + x;
                    if (mid(cells[ci]) != mid(pixels[p])) continue;
                    pixels[p] = cells[ci];
                    cells[ci] = air_id | (cells[ci] & 0xFFFF0000ull);
                    found = 1;
                }
                if (!found && wx >= 0 && wy >= 0 && wx < W && wy < H && is_air(&P, mid(cells[(size_t)wy * W + wx])))
                    pixels[p] = air_id;
            }
    }
    return 0;
}

/* ----------------------------------------------------------------------- */
/* Census                                                                   */
/* ----------------------------------------------------------------------- */
void orc_counts(const uint64_t *cells, size_t n, uint64_t counts[16]) {
    memset(counts, 0, 16 * sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) counts[cells[i] & 15]++;
}

/* Sum over cells of splitmix64(cell + index*GOLDEN): position-sensitive,
 * and the same in any summation order. */
uint64_t orc_checksum(const uint64_t *cells, size_t n) {
    uint64_t h = 0;
    for (size_t i = 0; i < n; i++) h += orc_splitmix64(cells[i] + (uint64_t)i * GOLDEN);
    return h;
}

/* ----------------------------------------------------------------------- */
/* Island check: world::physicsCheck (world.cpp:3330-3429)                  */
/* ----------------------------------------------------------------------- */
static int solid_cell(const fsg_config *cfg, uint64_t c) {
    const uint32_t m = (uint32_t)(c & 0xFF);
    return m < cfg->n_materials && cfg->materials[m].type == FSG_SOLID;
}

typedef struct { int x, y, dir; } flood_frame;

/* physicsCheck_flood (world.cpp:3407-3425): the recursive 4-way DFS, with an
 * explicit stack in the same visit order (x+1, y+1, x-1, y-1); a call
 * returns at once when more than 1000 cells were counted or out of bounds. */
static void flood(const uint64_t *cells, int W, int H, const fsg_config *cfg, int x0, int y0,
                  uint8_t *visited, int *count, int *minX, int *maxX, int *minY, int *maxY,
                  flood_frame *st) {
    static const int ddx[4] = {1, 0, -1, 0}, ddy[4] = {0, 1, 0, -1};
    int top = 0;
    st[0].x = x0, st[0].y = y0, st[0].dir = -1;
    while (top >= 0) {
        flood_frame *f = &st[top];
        if (f->dir < 0) { /* entering physicsCheck_flood(f->x, f->y) */
            const int x = f->x, y = f->y;
            if (*count > 1000 || x < 0 || x >= W || y < 0 || y >= H ||
                visited[(size_t)x + (size_t)y * W] || !solid_cell(cfg, cells[(size_t)x + (size_t)y * W])) {
                top--;
                continue;
            }
            if (x < *minX) *minX = x;
            if (x > *maxX) *maxX = x;
            if (y < *minY) *minY = y;
            if (y > *maxY) *maxY = y;
            visited[(size_t)x + (size_t)y * W] = 1;
            (*count)++;
            f->dir = 0;
        } else if (f->dir < 4) {
            const int d = f->dir++;
            top++;
            st[top].x = f->x + ddx[d], st[top].y = f->y + ddy[d], st[top].dir = -1;
        } else {
            top--;
        }
    }
}

static void physics_check_seq(uint64_t *cells, int W, int H, const fsg_config *cfg, fsg_island *seeds, int n,
                              uint64_t *pixels, size_t *needed, uint8_t *visited, flood_frame *st) {
    const uint64_t nothing = (uint64_t)cfg->mat_air; /* Tiles_NOTHING: colour 0, temperature 0 */
    *needed = 0;
    for (int i = 0; i < n; i++) {
        fsg_island *s = &seeds[i];
        const int x = s->x, y = s->y;
        s->status = FSG_ISLAND_NONE;
        s->count = 0;
        s->min_x = s->min_y = s->w = s->h = 0;
        s->pixel_offset = 0;
        if (x < 0 || y < 0 || x >= W || y >= H || !solid_cell(cfg, cells[(size_t)x + (size_t)y * W])) continue;
        memset(visited, 0, (size_t)W * H);
        int count = 0, minX = W, maxX = 0, minY = H, maxY = 0;
        flood(cells, W, H, cfg, x, y, visited, &count, &minX, &maxX, &minY, &maxY, st);
        if (count > 1000) {
            s->status = FSG_ISLAND_ATTACHED;
            continue;
        }
        if (count <= 0) continue;
        const int bw = maxX - minX + 1, bh = maxY - minY + 1;
        s->count = (uint32_t)count;
        s->min_x = minX, s->min_y = minY, s->w = bw, s->h = bh;
        s->status = count > 10 ? FSG_ISLAND_BODY : FSG_ISLAND_REMOVED;
        if (s->status == FSG_ISLAND_BODY) {
            s->pixel_offset = *needed;
            if (pixels)
                for (size_t k = 0; k < (size_t)bw * bh; k++) pixels[*needed + k] = nothing;
        }
        for (int yy = minY; yy <= maxY; yy++)
            for (int xx = minX; xx <= maxX; xx++) {
                const size_t j = (size_t)xx + (size_t)yy * W;
                if (!visited[j]) continue;
                if (s->status == FSG_ISLAND_BODY && pixels)
                    pixels[*needed + (size_t)(yy - minY) * bw + (xx - minX)] = cells[j];
                cells[j] = nothing;
            }
        if (s->status == FSG_ISLAND_BODY) *needed += (size_t)bw * bh;
    }
}

int orc_physics_check(uint64_t *cells, int W, int H, const fsg_config *cfg, fsg_island *seeds, int n,
                      uint64_t *pixels_out, size_t pixels_cap, size_t *pixels_needed) {
    if (!cells || !cfg || W <= 0 || H <= 0 || n < 0 || (n && !seeds)) return -1;
    uint8_t *visited = (uint8_t *)malloc((size_t)W * H);
    uint64_t *copy = (uint64_t *)malloc((size_t)W * H * 8);
    fsg_island *sc = (fsg_island *)malloc(sizeof(fsg_island) * (size_t)(n ? n : 1));
    flood_frame *st = (flood_frame *)malloc(sizeof(flood_frame) * 1100);
    if (!visited || !copy || !sc || !st) {
        free(visited), free(copy), free(sc), free(st);
        return -2;
    }
    /* dry run on a copy: the image cells needed */
    size_t needed = 0;
    memcpy(copy, cells, (size_t)W * H * 8);
    memcpy(sc, seeds, sizeof(fsg_island) * (size_t)n);
    physics_check_seq(copy, W, H, cfg, sc, n, NULL, &needed, visited, st);
    if (pixels_needed) *pixels_needed = needed;
    int rc = 0;
    if (needed && (!pixels_out || pixels_cap < needed)) rc = -1;
    else physics_check_seq(cells, W, H, cfg, seeds, n, pixels_out, &needed, visited, st);
    free(visited), free(copy), free(sc), free(st);
    return rc;
}

/* ----------------------------------------------------------------------- */
/* Texture pass (game.cpp:1994-2067)                                        */
/* ----------------------------------------------------------------------- */
/* The build's cell colour is rgba with bytes R,G,B,A in memory; the
 * reference keeps MaterialInstance::color = 0x00RRGGBB and Material::alpha
 * apart.  Each texel is assembled byte by byte as game.cpp does it, from
 * those reference-side values. */
static uint32_t texel_of_bytes(const uint8_t b[4]) {
    return (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24;
}

/* dpixels_ar for a non-AIR cell (game.cpp:2017-2021). */
static uint32_t cell_texel(uint64_t cell) {
    const uint32_t rgba = (uint32_t)(cell >> 32);
    const uint32_t color = (rgba & 0xFFu) << 16 | ((rgba >> 8) & 0xFFu) << 8 | ((rgba >> 16) & 0xFFu);
    const uint8_t alpha = (uint8_t)(rgba >> 24);
    uint8_t b[4];
    b[2] = (uint8_t)((color >> 0) & 0xff);
    b[1] = (uint8_t)((color >> 8) & 0xff);
    b[0] = (uint8_t)((color >> 16) & 0xff);
    b[3] = alpha;
    return texel_of_bytes(b);
}

/* dpixelsEmission_ar for a non-AIR cell (game.cpp:2023-2027). */
static uint32_t emit_texel(uint32_t emit) {
    uint8_t b[4];
    b[2] = (uint8_t)((emit >> 0) & 0xff);
    b[1] = (uint8_t)((emit >> 8) & 0xff);
    b[0] = (uint8_t)((emit >> 16) & 0xff);
    b[3] = (uint8_t)((emit >> 24) & 0xff);
    return texel_of_bytes(b);
}

int orc_render(const uint64_t *cells, const uint64_t *prev, int W, int H, const fsg_config *cfg,
               const fsg_render_params *p, uint32_t *lc, uint32_t *le, uint32_t *lf,
               fsg_render_stats *st) {
    if (!cells || !cfg || !p || W <= 0 || H <= 0) return -1;
    if (st) memset(st, 0, sizeof(*st));
    int64_t x0 = 0, y0 = 0, x1 = W, y1 = H;
    if (p->w > 0 && p->h > 0) {
        x0 = p->x0 < 0 ? 0 : p->x0;
        y0 = p->y0 < 0 ? 0 : p->y0;
        x1 = (int64_t)p->x0 + p->w < W ? (int64_t)p->x0 + p->w : W;
        y1 = (int64_t)p->y0 + p->h < H ? (int64_t)p->y0 + p->h : H;
    }
    for (int64_t y = y0; y < y1; y++)
        for (int64_t x = x0; x < x1; x++) {
            const size_t i = (size_t)y * (size_t)W + (size_t)x;
            const uint64_t c = cells[i];
            const uint32_t mat = (uint32_t)(c & 0xFF);
            const int air = mat < cfg->n_materials && cfg->materials[mat].type == FSG_AIR;
            if (prev && st && ((c ^ prev[i]) & 0xFFFFFFFF000000FFull)) {
                st->moved[mat & 15]++; /* movingTiles[mat->id]++ (game.cpp:1999) */
                st->had_dirty = 1;
            }
            if (air) { /* game.cpp:2000-2016 */
                if (lc && (p->layers & FSG_LAYER_CELLS)) lc[i] = 0;
                if (le && (p->layers & FSG_LAYER_EMISSION)) le[i] = 0;
                if (lf && (p->layers & FSG_LAYER_FIRE)) lf[i] = 0;
                continue;
            }
            if (lc && (p->layers & FSG_LAYER_CELLS)) lc[i] = cell_texel(c);
            if (le && (p->layers & FSG_LAYER_EMISSION)) le[i] = emit_texel(p->emit[mat & 15]);
            if (mat == p->fire_mat) { /* game.cpp:2029-2035 */
                if (lf && (p->layers & FSG_LAYER_FIRE)) lf[i] = cell_texel(c);
                if (st) st->fire_cells++;
            }
        }
    return 0;
}

/* ----------------------------------------------------------------------- */
/* The benchmark material table                                             */
/* ----------------------------------------------------------------------- */
/* Eight materials: the reference's GENERIC_AIR, STONE, GENERIC_SAND, WATER,
 * STEAM and LAVA (game_datastruct.cpp:72-110, densities 0, solid, 10, 1.5,
 * -1, 2 turned into ranks) plus oil (a liquid lighter than water) and wood
 * (a solid that burns), which the reference does not have. */
static void set_mat(fsg_material *m, const char *name, int type, int rank, int iters, int slip,
                    uint32_t rgba) {
    memset(m, 0, sizeof(*m));
    strncpy(m->name, name, sizeof(m->name) - 1);
    m->type = (uint8_t)type;
    m->rank = (uint8_t)rank;
    m->iterations = (uint8_t)iters;
    m->slipperyness = (uint8_t)slip;
    m->rgba = rgba;
}

void orc_default_config(fsg_config *cfg) {
    memset(cfg, 0, sizeof(*cfg));
    cfg->cell_iter = 3;
    cfg->reactions = 1;
    cfg->temp_every = 1;
    cfg->temp_phase = 0;
    cfg->chunk_skip = 0;
    cfg->wood_burn_temp = 300 * 16;
    cfg->mat_air = 0;
    cfg->mat_stone = 1;
    cfg->mat_water = 3;
    cfg->mat_steam = 5;
    cfg->mat_lava = 6;
    cfg->mat_wood = 7;
    cfg->n_materials = 8;
    set_mat(&cfg->materials[0], "air", FSG_AIR, 1, 0, 0, 0x00000000u);
    set_mat(&cfg->materials[1], "stone", FSG_SOLID, 14, 0, 0, 0xFF808080u);
    set_mat(&cfg->materials[2], "sand", FSG_SAND, 5, 2, 20, 0xFF50C8E6u);
    set_mat(&cfg->materials[3], "water", FSG_SOUP, 3, 6, 0, 0x80E06020u);
    set_mat(&cfg->materials[4], "oil", FSG_SOUP, 2, 3, 0, 0xE0104050u);
    set_mat(&cfg->materials[5], "steam", FSG_GAS, 0, 1, 0, 0xFF666666u);
    set_mat(&cfg->materials[6], "lava", FSG_SOUP, 4, 1, 0, 0xC00069FFu);
    set_mat(&cfg->materials[7], "wood", FSG_SOLID, 14, 0, 0, 0xFF2D5280u);
}

/* ------------------------------------------------------------------------ */
/* Chunk .pack images (chunk.cpp:74-330).                                    */
/* ------------------------------------------------------------------------ */

static void put_le(uint8_t *p, uint32_t v, int nbytes) {
    for (int k = 0; k < nbytes; k++) p[k] = (uint8_t)(v >> (8 * k));
}

static uint32_t get_le(const uint8_t *p, int nbytes) {
    uint32_t v = 0;
    for (int k = 0; k < nbytes; k++) v |= (uint32_t)p[k] << (8 * k);
    return v;
}

/* chunk.cpp:238-242: buf[i] = {(u16)mat->id, color, temperature} for tiles,
 * then layer2 at CHUNK_W * CHUNK_H + i.  The index is the material's id in
 * the file's material list (map->to_file), never the cell's MOVED bit; the
 * record's two trailing bytes are padding, which ChunkRead never reads, and
 * byte 10 carries MOVED in its bit 0. */
void orc_chunk_records(const uint64_t *tiles, size_t tstride, const uint64_t *layer2, size_t lstride,
                       const uint32_t *bg, size_t bstride, uint8_t *raw_a, uint8_t *raw_b,
                       const fsg_pack_map *map) {
    const int keep_moved = !map || !(map->flags & FSG_PACK_NO_MOVED);
    for (int layer = 0; layer < 2; layer++)
        for (int y = 0; y < 128; y++)
            for (int x = 0; x < 128; x++) {
                uint64_t c = 0;
                if (layer == 0) c = tiles[(size_t)y * tstride + x];
                else if (layer2) c = layer2[(size_t)y * lstride + x];
                const uint32_t id = (uint32_t)(c & 0xFFu);
                uint8_t *r = raw_a + ((size_t)layer * 16384 + (size_t)y * 128 + x) * 12;
                put_le(r, map ? map->to_file[id & 15u] : id, 4);   /* index */
                put_le(r + 4, (uint32_t)(c >> 32), 4);              /* color */
                put_le(r + 8, (uint32_t)((c >> 16) & 0xFFFFu), 2);  /* temperature */
                r[10] = (uint8_t)(keep_moved && (c & 0x100u) ? 1 : 0);
                r[11] = 0;
            }
    for (int y = 0; y < 128; y++)
        for (int x = 0; x < 128; x++) put_le(raw_b + ((size_t)y * 128 + x) * 4, bg ? bg[(size_t)y * bstride + x] : 0u, 4);
}

/* Reads an LZ4 length continuation: bytes are added while they are 255.
 * Fails when the input position passes lim after a byte (lz4.c:1916-1940). */
static int lz4_more(const uint8_t *src, int *ip, int lim, int *len, int cap) {
    int s;
    do {
        s = src[(*ip)++];
        *len += s;
        if (*ip > lim || *len > cap) return -1;
    } while (s == 255);
    return 0;
}

int orc_lz4_decode(const uint8_t *src, int n, uint8_t *dst, int cap) {
    if (n <= 0 || cap < 0) return -1;
    if (cap == 0) return (n == 1 && src[0] == 0) ? 0 : -1;
    int ip = 0, op = 0;
    while (ip < n) {
        const int token = src[ip++];
        int lit = token >> 4;
        if (lit == 15) {
            if (ip >= n - 15) return -1; /* initial check of the literal-length read */
            if (lz4_more(src, &ip, n - 15, &lit, cap)) return -1;
        }
        /* a literal run near either end must be the block's last sequence */
        if (op + lit > cap - 12 || ip + lit > n - 8) {
            if (ip + lit != n || op + lit > cap) return -1;
            memcpy(dst + op, src + ip, (size_t)lit);
            return op + lit;
        }
        memcpy(dst + op, src + ip, (size_t)lit);
        ip += lit;
        op += lit;
        const int offset = (int)get_le(src + ip, 2);
        ip += 2;
        int mlen = token & 15;
        if (mlen == 15 && lz4_more(src, &ip, n - 4, &mlen, cap)) return -1;
        mlen += 4;
        if (offset > op) return -1;         /* before the output start */
        if (op + mlen > cap - 5) return -1; /* the last 5 bytes are literals */
        /* offset 0: lz4.c zeroes the first word before copying the match onto
         * itself (lz4.c:2328-2334), so the match is zeros */
        for (int k = 0; k < mlen; k++, op++) dst[op] = offset ? dst[op - offset] : 0;
    }
    return -1;
}

static int lz4_put_len(uint8_t *dst, int op, int cap, int v) {
    if (v < 15) return op;
    for (v -= 15; v >= 255; v -= 255) {
        if (op >= cap) return -1;
        dst[op++] = 255;
    }
    if (op >= cap) return -1;
    dst[op++] = (uint8_t)v;
    return op;
}

/* One sequence: literals raw[ls, le), then (if mlen > 0) a match of mlen
 * bytes at offset off. */
static int lz4_put_seq(const uint8_t *raw, int ls, int le, int off, int mlen, uint8_t *dst, int op, int cap) {
    const int lit = le - ls;
    if (op >= cap) return -1;
    dst[op++] = (uint8_t)(((lit < 15 ? lit : 15) << 4) | (mlen ? (mlen - 4 < 15 ? mlen - 4 : 15) : 0));
    if ((op = lz4_put_len(dst, op, cap, lit)) < 0 || op + lit > cap) return -1;
    memcpy(dst + op, raw + ls, (size_t)lit);
    op += lit;
    if (!mlen) return op;
    if (op + 2 > cap) return -1;
    put_le(dst + op, (uint32_t)off, 2);
    return lz4_put_len(dst, op + 2, cap, mlen - 4);
}

int orc_lz4_encode_rows(const uint8_t *raw, int n, int rec, uint8_t *dst, int cap) {
    const int row = 128 * rec;
    if (rec <= 0 || n <= 12 || n % row) return -1;
    int op = 0, lit0 = 0; /* literals pending since lit0 */
    for (int r0 = 0; r0 < n; r0 += row) {
        int kind = -1, start = r0;
        for (int p = r0; p <= r0 + row; p++) {
            int k = -2; /* the row's end closes the last run */
            if (p < r0 + row) {
                const int eql = p >= rec && raw[p] == raw[p - rec];
                const int equ = p >= row && raw[p] == raw[p - row];
                k = p >= n - 12 ? 0 : eql ? 1 : equ ? 2 : 0;
            }
            if (k == kind) continue;
            if (kind > 0 && p - start >= 4) { /* a match run */
                op = lz4_put_seq(raw, lit0, start, kind == 1 ? rec : row, p - start, dst, op, cap);
                if (op < 0) return -1;
                lit0 = p;
            }
            kind = k;
            start = p;
        }
    }
    return lz4_put_seq(raw, lit0, n, 0, 0, dst, op, cap);
}

size_t orc_pack_chunk(const uint8_t *raw_a, const uint8_t *raw_b, int8_t phase, uint8_t *img) {
    const int ca = orc_lz4_encode_rows(raw_a, ORC_PACK_NA, 12, img + ORC_PACK_HDR, ORC_LZ4_BOUND(ORC_PACK_NA));
    const int cb = orc_lz4_encode_rows(raw_b, ORC_PACK_NB, 4, img + ORC_PACK_HDR + ca, ORC_LZ4_BOUND(ORC_PACK_NB));
    if (ca < 0 || cb < 0) return 0;
    img[0] = (uint8_t)phase; /* chunk.cpp:312-318 */
    put_le(img + 1, ORC_PACK_NA, 4);
    put_le(img + 5, (uint32_t)ca, 4);
    put_le(img + 9, ORC_PACK_NB, 4);
    put_le(img + 13, (uint32_t)cb, 4);
    return (size_t)ORC_PACK_HDR + (size_t)ca + (size_t)cb;
}

/* The material id a record index names: from_file, or the id itself below
 * n_materials; -1 for none (chunk.cpp:167-177 would index materials_array
 * out of range). */
static int file_material(const fsg_pack_map *map, uint32_t n_materials, uint32_t idx) {
    if (idx > 0xFFu) return -1;
    if (map) return map->from_file[idx];
    return idx < n_materials ? (int)idx : -1;
}

int orc_unpack_chunk(const uint8_t *img, size_t n, uint32_t n_materials, uint64_t *tiles, size_t tstride,
                     uint64_t *layer2, size_t lstride, uint32_t *bg, size_t bstride, int8_t *phase,
                     const fsg_pack_map *map) {
    const int keep_moved = !map || !(map->flags & FSG_PACK_NO_MOVED);
    if (n < ORC_PACK_HDR) return FSG_ERR_INVALID;
    if (phase) *phase = (int8_t)img[0];
    const int32_t sa = (int32_t)get_le(img + 1, 4), ca = (int32_t)get_le(img + 5, 4);
    const int32_t sb = (int32_t)get_le(img + 9, 4), cb = (int32_t)get_le(img + 13, 4);
    /* chunk.cpp:118-134: the sizes must be the chunk's */
    if (sa != ORC_PACK_NA || sb != ORC_PACK_NB || ca < 0 || cb < 0 ||
        (size_t)ORC_PACK_HDR + (size_t)ca + (size_t)cb > n)
        return FSG_ERR_INVALID;
    uint8_t *ra = (uint8_t *)malloc(ORC_PACK_NA), *rb = (uint8_t *)malloc(ORC_PACK_NB);
    int rc = FSG_ERR_INVALID;
    if (!ra || !rb) goto done;
    if (orc_lz4_decode(img + ORC_PACK_HDR, ca, ra, ORC_PACK_NA) != ORC_PACK_NA) goto done;
    if (orc_lz4_decode(img + ORC_PACK_HDR + ca, cb, rb, ORC_PACK_NB) != ORC_PACK_NB) goto done;
    for (int i = 0; i < 2 * 16384; i++) /* chunk.cpp:167-177: index -> materials_array */
        if (file_material(map, n_materials, get_le(ra + (size_t)i * 12, 4)) < 0) goto done;
    for (int i = 0; i < 2 * 16384; i++) {
        const uint8_t *r = ra + (size_t)i * 12;
        const uint64_t mat = (uint64_t)file_material(map, n_materials, get_le(r, 4)) |
                             (keep_moved && (r[10] & 1u) ? 0x100u : 0u);
        const uint64_t c = mat | ((uint64_t)get_le(r + 8, 2) << 16) | ((uint64_t)get_le(r + 4, 4) << 32);
        const int y = (i >> 7) & 127, x = i & 127;
        if (i < 16384) tiles[(size_t)y * tstride + x] = c;
        else if (layer2) layer2[(size_t)y * lstride + x] = c;
    }
    if (bg)
        for (int i = 0; i < 16384; i++) bg[(size_t)(i >> 7) * bstride + (i & 127)] = get_le(rb + (size_t)i * 4, 4);
    rc = FSG_OK;
done:
    free(ra);
    free(rb);
    return rc;
}
