/*
 * refseq.cpp -- restatement of the reference's sequential world::tick.
 *
 * TEST INFRASTRUCTURE ONLY (see fsg_oracle.h): the timed CPU baseline and
 * the comparator for conservation and settled column-height statistics.
 * Never linked into the product.
 *
 * It keeps the reference's schedule (source/engine/world.cpp:1036-1948):
 * cell_iter iterations x 4 phases of a 2x2 checkerboard of 128x128 chunks
 * (1057-1074), one task per chunk on a thread pool (1077; the reference uses
 * 16 threads, world.cpp:59), each chunk scanned bottom-to-top, left-to-right
 * in three passes A/B/C (1084, 1589, 1823) with a per-phase tickVisited mask
 * (1066-1067).  Its rules are the reference's, with the benchmark material
 * table:
 *   - densities are the material ranks;
 *   - SOUP uses the reference's discrete rule (the commented-out code of
 *     world.cpp:1541-1567, 1750-1798, 1839-1861) instead of the float
 *     liquid-mass model, so cell counts are conserved;
 *   - free-fall particles (CellData, 1218-1225) stay on the grid as plain
 *     swaps and physicsCheck (1929-1934) is off;
 *   - declared reactions only: water above lava, wood above its burn
 *     temperature; steam condensation (1883-1888) is excluded;
 *   - rand() becomes a splitmix64 stream per chunk task, so a run is
 *     reproducible for any thread count;
 *   - neighbour reads use getTile's out-of-world = solid rule (999-1002)
 *     instead of the reference's 128-cell tickZone margin.
 * Temperature uses the same integer 5-point step as the det mode.
 */
#include "fsg_oracle.h"
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int CH = 128; /* CHUNK_W/CHUNK_H, core/const.h:19-20 */
constexpr uint64_t WALL = 15;

struct Props {
    uint8_t type[16], rank[16], iters[16], slip[16];
    int16_t add_temp[16];
    uint32_t rgba[16];
};

struct Rng {
    uint64_t st;
    uint32_t next() { return (uint32_t)(orc_splitmix64(st++) >> 32); }
    int operator()(int n) { return (int)(next() % (uint32_t)n); } /* rand() % n */
};

struct World {
    uint64_t *c;
    int W, H;
    const fsg_config *cfg;
    Props P;
    uint8_t *visited;

    static int mid(uint64_t v) { return (int)(v & 0xFF); }
    static int temp(uint64_t v) { return (int)(int16_t)(uint16_t)(v >> 16); }
    bool in(int x, int y) const { return x >= 0 && y >= 0 && x < W && y < H; }
    uint64_t get(int x, int y) const { return in(x, y) ? c[(size_t)y * W + x] : WALL; }
    void set(int x, int y, uint64_t v) {
        if (in(x, y)) c[(size_t)y * W + x] = v;
    }
    void visit(int x, int y) {
        if (in(x, y)) visited[(size_t)y * W + x] = 1;
    }
    int type(uint64_t v) const { return P.type[mid(v)]; }
    /* canMoveBelow form (world.cpp:1206): AIR, or non-SOLID and lighter. */
    bool can_enter(uint64_t self, uint64_t other) const {
        const int t = type(other);
        return t == FSG_AIR || (t != FSG_SOLID && P.rank[mid(other)] < P.rank[mid(self)]);
    }
    uint64_t transform(uint64_t v, int m) const {
        return (uint64_t)(uint32_t)m | (v & 0xFFFF0000ull) | ((uint64_t)P.rgba[m] << 32);
    }

    void pass_a(int x, int y, int iter, Rng &rnd) {
        const size_t idx = (size_t)y * W + x;
        uint64_t tile = c[idx];
        const int m = mid(tile);
        if (cfg->reactions && m == cfg->mat_wood && temp(tile) > cfg->wood_burn_temp) {
            c[idx] = transform(tile, cfg->mat_steam);
            visited[idx] = 1;
            return;
        }
        if (iter >= P.iters[m]) { /* world.cpp:1093-1096 */
            visited[idx] = 1;
            return;
        }
        const int t = P.type[m];
        if (t == FSG_SAND) { /* world.cpp:1148-1268 */
            const uint64_t below = get(x, y + 1);
            if (!can_enter(tile, below)) return;
            const bool cl = can_enter(tile, get(x - 1, y + 1));
            const bool cr = can_enter(tile, get(x + 1, y + 1));
            if (!((cl || cr) && rnd(20) == 0)) {
                c[idx] = below;
                if (rnd(2) == 0) tile |= FSG_MOVED;
                set(x, y + 1, tile);
                visit(x, y + 1);
                if (rnd(2) == 0) {
                    if (x > 0 && type(get(x - 1, y + 1)) == FSG_SAND && rnd(2) == 0)
                        set(x - 1, y + 1, get(x - 1, y + 1) | FSG_MOVED);
                    if (x < W - 1 && type(get(x + 1, y + 1)) == FSG_SAND && rnd(2) == 0)
                        set(x + 1, y + 1, get(x + 1, y + 1) | FSG_MOVED);
                }
            }
        } else if (t == FSG_SOUP) {
            const uint64_t below = get(x, y + 1);
            if (cfg->reactions && m == cfg->mat_water && mid(below) == cfg->mat_lava) {
                /* world.cpp:1519-1537 */
                c[idx] = transform(tile, cfg->mat_steam);
                set(x, y + 1, transform(below, cfg->mat_stone));
                visit(x, y + 1);
                for (int xx = -1; xx <= 1; xx++)
                    for (int yy = 0; yy <= 2; yy++)
                        if (in(x + xx, y + yy) && mid(get(x + xx, y + yy)) == cfg->mat_lava) {
                            set(x + xx, y + yy, transform(get(x + xx, y + yy), cfg->mat_stone));
                            visit(x + xx, y + yy);
                        }
                return;
            }
            /* discrete fall (commented template, world.cpp:1541-1567) */
            if (!can_enter(tile, below)) return;
            const bool cl = can_enter(tile, get(x - 1, y + 1));
            const bool cr = can_enter(tile, get(x + 1, y + 1));
            if (!((cl || cr) && rnd(10) == 0)) {
                c[idx] = below;
                set(x, y + 1, tile);
                visit(x, y + 1);
            }
        } else if (t == FSG_GAS) { /* world.cpp:1569-1585 */
            const int above = type(get(x, y - 1));
            const int al = type(get(x - 1, y - 1)), ar = type(get(x + 1, y - 1));
            if (above == FSG_AIR && !((al == FSG_AIR || ar == FSG_AIR) && rnd(2) == 0)) {
                c[idx] = get(x, y - 1);
                set(x, y - 1, tile);
                visit(x, y - 1);
            }
        }
    }

    void pass_b(int x, int y, Rng &rnd) {
        const size_t idx = (size_t)y * W + x;
        uint64_t tile = c[idx];
        const int m = mid(tile);
        const int t = P.type[m];
        if (t == FSG_SAND) { /* world.cpp:1602-1727 */
            const uint64_t bl = get(x - 1, y + 1), br = get(x + 1, y + 1);
            const bool cl = can_enter(tile, bl), cr = can_enter(tile, br);
            bool stopped = !(tile & FSG_MOVED);
            const int slip = P.slip[m] ? P.slip[m] : 1;
            if (stopped) {
                int drop = 0;
                for (int pil = 0; pil < 10; pil++) {
                    const int l = type(get(x - 1, y + 1 + pil)), r = type(get(x + 1, y + 1 + pil));
                    if (l == FSG_AIR || r == FSG_AIR) drop++;
                }
                const int max_stab = (int)(8 / std::sqrt((double)slip) + 1);
                if (drop + 1 - max_stab > 0) {
                    const int chance = 1000 / (drop + 1 - max_stab);
                    if (chance < 1000 && rnd(chance) == 0) {
                        stopped = false;
                        c[idx] |= FSG_MOVED;
                        tile |= FSG_MOVED;
                    }
                }
            }
            if (stopped || !(cl || cr)) {
                c[idx] &= ~(uint64_t)FSG_MOVED;
                return;
            }
            const bool should = rnd(2 * slip) != 0;
            if (should && (cl || cr) && rnd(2) == 0) {
                if (type(get(x, y + 1)) == FSG_SAND && rnd(2) == 0)
                    set(x, y + 1, get(x, y + 1) | FSG_MOVED);
            }
            if (should && cl && (!cr || rnd(2) == 0)) {
                if (type(get(x - 1, y)) == FSG_AIR) {
                    const uint64_t side = get(x - 1, y);
                    set(x - 1, y, bl);
                    visit(x - 1, y);
                    c[idx] = side;
                } else {
                    c[idx] = bl;
                    visited[idx] = 1;
                }
                if (rnd(20 * slip) == 0) tile &= ~(uint64_t)FSG_MOVED;
                set(x - 1, y + 1, tile);
                visit(x - 1, y + 1);
            } else if (should && cr) {
                if (type(get(x + 1, y)) == FSG_AIR) {
                    const uint64_t side = get(x + 1, y);
                    set(x + 1, y, br);
                    c[idx] = side;
                } else {
                    c[idx] = br;
                    visited[idx] = 1;
                }
                if (rnd(20 * slip) == 0) tile &= ~(uint64_t)FSG_MOVED;
                set(x + 1, y + 1, tile);
                visit(x + 1, y + 1);
            } else {
                c[idx] &= ~(uint64_t)FSG_MOVED;
            }
        } else if (t == FSG_SOUP) { /* commented template, world.cpp:1750-1798 */
            const uint64_t bl = get(x - 1, y + 1), br = get(x + 1, y + 1);
            const bool cl = can_enter(tile, bl), cr = can_enter(tile, br);
            if (!(cl || cr)) return;
            const bool ml = can_enter(tile, get(x - 1, y)), mr = can_enter(tile, get(x + 1, y));
            if (!((ml || mr) && rnd(10) == 0)) {
                if (cl && !(cr && rnd(2) == 0)) {
                    if (type(get(x - 1, y)) == FSG_AIR) {
                        const uint64_t side = get(x - 1, y);
                        set(x - 1, y, bl);
                        c[idx] = side;
                    } else {
                        c[idx] = bl;
                    }
                    set(x - 1, y + 1, tile);
                    visit(x - 1, y + 1);
                } else if (cr) {
                    if (type(get(x + 1, y)) == FSG_AIR) {
                        const uint64_t side = get(x + 1, y);
                        set(x + 1, y, br);
                        c[idx] = side;
                    } else {
                        c[idx] = br;
                    }
                    set(x + 1, y + 1, tile);
                    visit(x + 1, y + 1);
                }
            }
        } else if (t == FSG_GAS) { /* world.cpp:1799-1819 */
            const int al = type(get(x - 1, y - 1)), ar = type(get(x + 1, y - 1));
            if (al == FSG_AIR && !(ar == FSG_AIR && rnd(2) == 0)) {
                c[idx] = get(x - 1, y - 1);
                set(x - 1, y - 1, tile);
                visit(x - 1, y - 1);
            } else if (ar == FSG_AIR) {
                c[idx] = get(x + 1, y - 1);
                set(x + 1, y - 1, tile);
                visit(x + 1, y - 1);
            }
        }
    }

    void pass_c(int x, int y, Rng &rnd) {
        const size_t idx = (size_t)y * W + x;
        const uint64_t tile = c[idx];
        const int t = type(tile);
        if (t == FSG_SOUP) { /* commented template, world.cpp:1839-1861; side chosen
                                by a fair coin instead of the template's
                                always-left (rand() % 2 == 5 never holds) */
            const uint64_t l = get(x - 1, y), r = get(x + 1, y);
            const bool ml = can_enter(tile, l), mr = can_enter(tile, r);
            if (ml && !(mr && rnd(2) == 0)) {
                c[idx] = l;
                set(x - 1, y, tile);
                visit(x - 1, y);
            } else if (mr) {
                c[idx] = r;
                set(x + 1, y, tile);
                visit(x + 1, y);
            }
        } else if (t == FSG_GAS) { /* world.cpp:1862-1882 */
            const int l = type(get(x - 1, y)), r = type(get(x + 1, y));
            if (l == FSG_AIR && !(r == FSG_AIR && rnd(2) == 0)) {
                c[idx] = get(x - 1, y);
                set(x - 1, y, tile);
                visit(x - 1, y);
            } else if (r == FSG_AIR) {
                c[idx] = get(x + 1, y);
                set(x + 1, y, tile);
                visit(x + 1, y);
            }
        }
    }

    void chunk(int cx, int cy, int iter, Rng &rnd) {
        for (int pass = 0; pass < 3; pass++) {
            for (int dy = CH - 1; dy >= 0; dy--) {
                const int y = cy + dy;
                if (y < 0 || y >= H) continue;
                for (int dx = 0; dx < CH; dx++) {
                    const int x = cx + dx;
                    if (x < 0 || x >= W) continue;
                    if (visited[(size_t)y * W + x]) continue;
                    if (pass == 0) pass_a(x, y, iter, rnd);
                    else if (pass == 1) pass_b(x, y, rnd);
                    else pass_c(x, y, rnd);
                }
            }
        }
    }
};

void temperature(World &w) {
    std::vector<int16_t> nt((size_t)w.W * w.H);
    for (int y = 0; y < w.H; y++)
        for (int x = 0; x < w.W; x++) {
            const uint64_t c = w.c[(size_t)y * w.W + x];
            const int T = World::temp(c);
            const int tu = y > 0 ? World::temp(w.c[(size_t)(y - 1) * w.W + x]) : T;
            const int td = y < w.H - 1 ? World::temp(w.c[(size_t)(y + 1) * w.W + x]) : T;
            const int tl = x > 0 ? World::temp(w.c[(size_t)y * w.W + x - 1]) : T;
            const int tr = x < w.W - 1 ? World::temp(w.c[(size_t)y * w.W + x + 1]) : T;
            int v = T + ((tu + td + tl + tr - 4 * T) >> 3) + w.P.add_temp[World::mid(c)];
            v = v > 32767 ? 32767 : (v < -32768 ? -32768 : v);
            nt[(size_t)y * w.W + x] = (int16_t)v;
        }
    for (size_t i = 0; i < nt.size(); i++)
        w.c[i] = (w.c[i] & 0xFFFFFFFF0000FFFFull) | ((uint64_t)(uint16_t)nt[i] << 16);
}

} // namespace

extern "C" int refseq_run(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed,
                          int nticks, int threads) {
    if (!cells || W <= 0 || H <= 0 || !cfg) return FSG_ERR_INVALID;
    World w;
    w.c = cells;
    w.W = W;
    w.H = H;
    w.cfg = cfg;
    for (int i = 0; i < 16; i++) {
        w.P.type[i] = FSG_SOLID;
        w.P.rank[i] = 15;
        w.P.iters[i] = 0;
        w.P.slip[i] = 0;
        w.P.add_temp[i] = 0;
        w.P.rgba[i] = 0;
    }
    for (uint32_t i = 0; i < cfg->n_materials && i < FSG_MAX_MATERIALS; i++) {
        w.P.type[i] = cfg->materials[i].type;
        w.P.rank[i] = cfg->materials[i].rank;
        w.P.iters[i] = cfg->materials[i].iterations;
        w.P.slip[i] = cfg->materials[i].slipperyness;
        w.P.add_temp[i] = cfg->materials[i].add_temp;
        w.P.rgba[i] = cfg->materials[i].rgba;
    }
    std::vector<uint8_t> visited((size_t)W * H);
    w.visited = visited.data();
    if (threads < 1) threads = 1;

    /* Persistent pool (the reference keeps WorldSystem::tickPool alive,
     * utility.hpp:308-470): workers wait for a phase, drain its chunk list,
     * and the main thread joins on a counter like the futures of
     * world.cpp:1903-1911. */
    std::vector<std::pair<int, int>> chunks;
    std::atomic<size_t> next{0};
    std::atomic<int> done{0};
    std::mutex mu;
    std::condition_variable cv_start, cv_done;
    uint64_t generation = 0;
    bool quit = false;
    int cur_tick = 0, cur_iter = 0, cur_tk = 0;
    auto drain = [&]() {
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= chunks.size()) break;
            Rng rnd;
            rnd.st = orc_splitmix64(seed ^ orc_splitmix64(
                ((uint64_t)cur_tick << 32) ^ ((uint64_t)cur_iter << 28) ^ ((uint64_t)cur_tk << 24) ^
                ((uint64_t)(chunks[k].first / CH) << 12) ^ (uint64_t)(chunks[k].second / CH)));
            w.chunk(chunks[k].first, chunks[k].second, cur_iter, rnd);
        }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < threads; i++) {
        pool.emplace_back([&]() {
            uint64_t seen = 0;
            for (;;) {
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv_start.wait(lk, [&] { return quit || generation != seen; });
                    if (quit) return;
                    seen = generation;
                }
                drain();
                if (done.fetch_add(1) + 1 == threads - 1) {
                    std::lock_guard<std::mutex> lk(mu);
                    cv_done.notify_one();
                }
            }
        });
    }
    for (int tick = 0; tick < nticks; tick++) {
        for (int iter = 0; iter < (int)cfg->cell_iter; iter++) {
            for (int tk = 0; tk < 4; tk++) {
                const int ofx = tk % 2, ofy = 1 - (tk % 4) / 2; /* world.cpp:1057-1060 */
                std::memset(visited.data(), 0, visited.size());
                chunks.clear();
                for (int cx = ofx * CH; cx < W; cx += 2 * CH)
                    for (int cy = ofy * CH; cy < H; cy += 2 * CH) chunks.emplace_back(cx, cy);
                cur_tick = tick;
                cur_iter = iter;
                cur_tk = tk;
                next.store(0);
                done.store(0);
                if (threads > 1) {
                    std::lock_guard<std::mutex> lk(mu);
                    generation++;
                    cv_start.notify_all();
                }
                drain();
                if (threads > 1) {
                    std::unique_lock<std::mutex> lk(mu);
                    cv_done.wait(lk, [&] { return done.load() == threads - 1; });
                }
            }
        }
        if (cfg->temp_every > 0 && ((uint64_t)tick % cfg->temp_every) == cfg->temp_phase)
            temperature(w);
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        quit = true;
        cv_start.notify_all();
    }
    for (auto &th : pool) th.join();
    return 0;
}
