// fsg_host.cpp -- host-only parts of the C-ABI: status strings, the default
// material table, the counter-RNG key schedule and the world generator.
//
// The generator is not on the device path; it builds the synthetic worlds of
// the benchmark (terrain from 1-D value noise, discs, a water basin, or the
// striped all-active world) in host memory for fsg_world_upload.
#include "fsg.h"

#include <cstring>
#include <vector>

#include "fsg_generate.hpp"

#define GOLDEN 0x9E3779B97F4A7C15ull

extern "C" const char *fsg_status_string(int st) {
    switch (st) {
    case FSG_OK: return "ok";
    case FSG_ERR_INVALID: return "invalid argument";
    case FSG_ERR_NOMEM: return "out of memory";
    case FSG_ERR_DEVICE: return "device error";
    case FSG_ERR_BUDGET: return "HBM state budget exceeded";
    case FSG_ERR_STATE: return "invalid state";
    case FSG_ERR_CHECKSUM: return "state checksum mismatch";
    case FSG_ERR_TIMEOUT: return "deadline passed";
    case FSG_ERR_CANCELLED: return "cancelled";
    case FSG_ERR_NOTFOUND: return "not found";
    case FSG_ERR_AGAIN: return "not ready";
    default: return "unknown status";
    }
}

extern "C" uint64_t fsg_splitmix64(uint64_t x) {
    uint64_t z = x + GOLDEN;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" uint64_t fsg_phase_key(uint64_t seed, uint64_t g) {
    return fsg_splitmix64(fsg_splitmix64(seed) ^ g);
}

namespace {
struct MatDef {
    const char *name;
    uint8_t type, rank, iters, slip;
    uint32_t rgba;
};
// Reference materials (game_datastruct.cpp:72-110): GENERIC_AIR, STONE,
// GENERIC_SAND (slip 20, 2 iterations), WATER (6 iterations), STEAM,
// LAVA (1 iteration); densities become ranks: steam < air < oil < water <
// lava < sand.  Oil and wood are additions of the benchmark worlds.
const MatDef kMats[8] = {
    {"air", FSG_AIR, 1, 0, 0, 0x00000000u},   {"stone", FSG_SOLID, 14, 0, 0, 0xFF808080u},
    {"sand", FSG_SAND, 5, 2, 20, 0xFF50C8E6u}, {"water", FSG_SOUP, 3, 6, 0, 0x80E06020u},
    {"oil", FSG_SOUP, 2, 3, 0, 0xE0104050u},  {"steam", FSG_GAS, 0, 1, 0, 0xFF666666u},
    {"lava", FSG_SOUP, 4, 1, 0, 0xC00069FFu}, {"wood", FSG_SOLID, 14, 0, 0, 0xFF2D5280u},
};
enum { M_AIR = 0, M_STONE = 1, M_SAND = 2, M_WATER = 3, M_OIL = 4, M_STEAM = 5, M_LAVA = 6, M_WOOD = 7 };
} // namespace

extern "C" int fsg_default_config(fsg_config *cfg) {
    if (!cfg) return FSG_ERR_INVALID;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->cell_iter = 3;          // global.lua:47
    cfg->reactions = 1;
    cfg->temp_every = 1;
    cfg->temp_phase = 0;
    cfg->chunk_skip = 0;
    cfg->wood_burn_temp = 300 * 16;
    cfg->mat_air = M_AIR;
    cfg->mat_water = M_WATER;
    cfg->mat_lava = M_LAVA;
    cfg->mat_steam = M_STEAM;
    cfg->mat_stone = M_STONE;
    cfg->mat_wood = M_WOOD;
    cfg->n_materials = 8;
    for (int i = 0; i < 8; i++) {
        fsg_material &m = cfg->materials[i];
        std::strncpy(m.name, kMats[i].name, sizeof(m.name) - 1);
        m.type = kMats[i].type;
        m.rank = kMats[i].rank;
        m.iterations = kMats[i].iters;
        m.slipperyness = kMats[i].slip;
        m.add_temp = 0;
        m.rgba = kMats[i].rgba;
    }
    return FSG_OK;
}

// ---------------------------------------------------------------------------
// Generator
// ---------------------------------------------------------------------------
namespace {

struct Stream { // splitmix64 sequence
    uint64_t st;
    uint64_t next() {
        uint64_t v = fsg_splitmix64(st);
        st += GOLDEN;
        return v;
    }
    uint32_t below(uint32_t n) { return (uint32_t)((next() >> 32) % n); }
};

constexpr int16_t kAmbient = 20 * 16;  // 20 degrees
constexpr int16_t kLavaTemp = 1200 * 16;

inline uint32_t jitter(uint32_t rgba, uint64_t h) {
    // +-8 per colour channel, alpha untouched
    uint32_t out = rgba & 0xFF000000u;
    for (int ch = 0; ch < 3; ch++) {
        int v = (int)((rgba >> (8 * ch)) & 0xFF) + (int)((h >> (8 * ch)) & 15) - 8;
        v = v < 0 ? 0 : (v > 255 ? 255 : v);
        out |= (uint32_t)v << (8 * ch);
    }
    return out;
}

inline fsg_cell make(int m, int x, int y, uint64_t seed) {
    fsg_cell c;
    c.mat = (uint16_t)m;
    c.temp = m == M_LAVA ? kLavaTemp : kAmbient;
    const uint64_t h = fsg_splitmix64(seed ^ ((uint64_t)(uint32_t)x << 20) ^ ((uint64_t)(uint32_t)y << 42) ^ 0xC0105ull);
    c.rgba = m == M_AIR ? 0u : jitter(kMats[m].rgba, h);
    return c;
}

// 1-D value noise, 4 octaves, integer: lattice values are the top 16 bits of
// splitmix64(seed, octave, lattice index); linear interpolation; octave
// weights 8:4:2:1.  Returns [0, 65535].
int value_noise(uint64_t seed, int x, int W) {
    int acc = 0;
    const int wsum = 15;
    for (int o = 0; o < 4; o++) {
        int period = (W / 4) >> o;
        if (period < 2) period = 2;
        const int i = x / period, f = x % period;
        auto lat = [&](int k) {
            return (int)(fsg_splitmix64(seed ^ ((uint64_t)(o + 1) << 56) ^ (uint64_t)(uint32_t)k) >> 48);
        };
        const int v0 = lat(i), v1 = lat(i + 1);
        const int v = v0 + (int)(((int64_t)(v1 - v0) * f) / period);
        acc += (8 >> o) * v;
    }
    return acc / wsum;
}

void disc(fsg_cell *g, int W, int H, int cx, int cy, int rad, int m, uint64_t seed) {
    for (int y = cy - rad; y <= cy + rad; y++) {
        if (y < 0 || y >= H) continue;
        for (int x = cx - rad; x <= cx + rad; x++) {
            if (x < 0 || x >= W) continue;
            const int dx = x - cx, dy = y - cy;
            if (dx * dx + dy * dy > rad * rad) continue;
            fsg_cell &c = g[(size_t)y * W + x];
            if (c.mat == M_STONE) continue;  // discs do not carve the terrain
            c = make(m, x, y, seed);
        }
    }
}

void basin_discs(int W, int H, uint64_t seed, int n_sand, int n_lava, int n_oil, std::vector<fsg_gen_disc> &out) {
    out.clear();
    Stream rs{fsg_splitmix64(seed ^ 0x5A4Dull)};
    for (int i = 0; i < n_sand; i++) {
        const int cx = (int)rs.below((uint32_t)W), cy = (int)rs.below((uint32_t)H);
        const int r = 4 + (int)rs.below(17);
        out.push_back({cx, cy, r, M_SAND});
    }
    Stream rl{fsg_splitmix64(seed ^ 0x1A7Aull)};
    for (int i = 0; i < n_lava; i++) {
        const int cx = (int)rl.below((uint32_t)W), cy = (int)rl.below((uint32_t)H);
        const int r = 8 + (int)rl.below(25);
        out.push_back({cx, cy, r, M_LAVA});
    }
    Stream ro{fsg_splitmix64(seed ^ 0x0111ull)};
    for (int i = 0; i < n_oil; i++) {
        const int cx = (int)ro.below((uint32_t)W), cy = (int)ro.below((uint32_t)H);
        const int r = 4 + (int)ro.below(17);
        out.push_back({cx, cy, r, M_OIL});
    }
}

void gen_basin(fsg_cell *g, int W, int H, uint64_t seed, int n_sand, int n_lava,
               int n_oil) {
    // terrain: stone below H - noise*0.4*H
    for (int x = 0; x < W; x++) {
        const int n = value_noise(seed, x, W);
        const int h = (int)(((int64_t)n * 2 * H) / (5 * 65536));
        for (int y = 0; y < H; y++)
            g[(size_t)y * W + x] = make(y >= H - h ? M_STONE : M_AIR, x, y, seed);
    }
    // water fills the basin up to 0.55*H above the bottom
    const int level = H - (H * 55) / 100;
    for (int y = level; y < H; y++)
        for (int x = 0; x < W; x++)
            if (g[(size_t)y * W + x].mat == M_AIR) g[(size_t)y * W + x] = make(M_WATER, x, y, seed);
    std::vector<fsg_gen_disc> discs;
    basin_discs(W, H, seed, n_sand, n_lava, n_oil, discs);
    for (const fsg_gen_disc &d : discs) disc(g, W, H, d.cx, d.cy, d.r, d.m, seed);
}

// Disc counts of a basin kind (1, 2, 4); false for other kinds.
bool basin_counts(int kind, int W, int H, int &n_sand, int &n_lava, int &n_oil) {
    switch (kind) {
    case 1: n_sand = 64, n_lava = 0, n_oil = 0; return true;
    case 2: n_sand = 64, n_lava = 256, n_oil = 128; return true;
    case 4: {
        // kind 2 mix scaled to the world's area relative to 4096^2
        const double a = (double)W * (double)H / (4096.0 * 4096.0);
        auto sc = [&](int n) { int v = (int)(n * a + 0.5); return v < 1 ? 1 : v; };
        n_sand = sc(64), n_lava = sc(256), n_oil = sc(128);
        return true;
    }
    default: return false;
    }
}

// All-active world: a water band (rows [H/10, 2H/10)) and a steam band
// (rows [8H/10, 9H/10)), each 10% of rows; every other row alternates
// between 50% sand on a random per-cell mask and air; bottom row stone.
void gen_striped(fsg_cell *g, int W, int H, uint64_t seed) {
    for (int y = 0; y < H; y++) {
        int band = -1;
        if (y * 10 >= H && y * 10 < 2 * H) band = M_WATER;
        if (y * 10 >= 8 * H && y * 10 < 9 * H) band = M_STEAM;
        for (int x = 0; x < W; x++) {
            int m;
            if (y == H - 1) m = M_STONE;
            else if (band >= 0) m = band;
            else if ((y & 1) == 0)
                m = (fsg_splitmix64(seed ^ ((uint64_t)(uint32_t)y << 32) ^ (uint64_t)(uint32_t)x) >> 63) ? M_SAND : M_AIR;
            else m = M_AIR;
            fsg_cell c;
            c.mat = (uint16_t)m;
            c.temp = kAmbient;
            c.rgba = m == M_AIR ? 0u : kMats[m].rgba;
            g[(size_t)y * W + x] = c;
        }
    }
}

} // namespace

extern "C" int fsg_generate(int kind, int W, int H, uint64_t seed, fsg_cell *out) {
    if (!out || W <= 0 || H <= 0 || W > FSG_MAX_SIDE || H > FSG_MAX_SIDE) return FSG_ERR_INVALID;
    if (kind == 3) {
        gen_striped(out, W, H, seed);
        return FSG_OK;
    }
    int ns, nl, no;
    if (!basin_counts(kind, W, H, ns, nl, no)) return FSG_ERR_INVALID;
    gen_basin(out, W, H, seed, ns, nl, no);
    return FSG_OK;
}

// The generator's plan for the device path (fsg_world_generate): the disc
// list of a basin kind in drawing order and the constants of make().
int fsg_gen_plan(int kind, int W, int H, uint64_t seed, std::vector<fsg_gen_disc> &discs, uint32_t rgba[8],
                 int16_t &ambient, int16_t &lava_temp) {
    for (int i = 0; i < 8; i++) rgba[i] = kMats[i].rgba;
    ambient = kAmbient;
    lava_temp = kLavaTemp;
    discs.clear();
    if (kind == 3) return FSG_OK;
    int ns, nl, no;
    if (!basin_counts(kind, W, H, ns, nl, no)) return FSG_ERR_INVALID;
    basin_discs(W, H, seed, ns, nl, no, discs);
    return FSG_OK;
}
