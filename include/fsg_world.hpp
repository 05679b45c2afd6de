// fsg_world.hpp -- C++ facade over the C-ABI with the member names of the
// reference's ME::world (source/engine/world.hpp:63-193), so the callers in
// game::tick (game.cpp:1659-2201) keep their shape:
//
//   world.init(w, h, seed)            world::init (world.cpp:43-172)
//   world.tick()                      world::tick (world.cpp:1036-1948)
//   world.tickTemperature()           world::tickTemperature (world.cpp:1950-2004)
//   world.getTile(x, y)               world::getTile (world.cpp:999-1002)
//   world.setTile(x, y, cell)         world::setTile (world.cpp:1004-1008)
//   world.stampBodies(...)            the stamp loop of game.cpp:1760-1815
//   world.unstampBodies(...)          the unstamp loop of game.cpp:1896-1983
//   world.updateTextures(...)         the texture pass of game.cpp:1994-2067
//   world.writeChunks(...)            Chunk::ChunkWrite (chunk.cpp:219-330) per chunk
//   world.readChunks(...)             Chunk::ChunkRead (chunk.cpp:74-217) per chunk
//
// Error behaviour follows the reference: getTile out of bounds returns the
// solid wall cell, setTile out of bounds is ignored.  Device failures, which
// the reference cannot have, throw fsg::error from the facade (the C-ABI
// itself never throws).
//
// getTile reads a host mirror of the grid in 64 x 64 pages: a page is
// downloaded on its first read (or by region / prefetch) and stays valid
// until a call of this facade changes the grid on the device (tick, stamp,
// load, readChunks, ...).  setTile writes through to a mirrored page and
// queues the edit for the device, so a getTile after a setTile sees the new
// cell as in the reference.  A caller that changes the grid through
// handle() calls invalidate().
#ifndef FSG_WORLD_HPP
#define FSG_WORLD_HPP

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "fsg.h"

namespace fsg {

struct error : std::runtime_error {
    int status;
    error(int st, const char *what) : std::runtime_error(std::string(what) + ": " + fsg_status_string(st)), status(st) {}
};

inline void check(int st, const char *what) {
    if (st != FSG_OK) throw error(st, what);
}

// The cell type (MaterialInstance in the reference).
using MaterialInstance = fsg_cell;

class context {
  public:
    explicit context(int device = 0, uint64_t hbm_budget = 0) {
        fsg_ctx_config cfg{device, hbm_budget};
        check(fsg_ctx_create(&cfg, &ctx_), "fsg_ctx_create");
    }
    ~context() {
        if (ctx_) fsg_ctx_destroy(ctx_);
    }
    context(const context &) = delete;
    context &operator=(const context &) = delete;
    fsg_ctx *get() const { return ctx_; }

  private:
    fsg_ctx *ctx_ = nullptr;
};

class world {
  public:
    int width = 0, height = 0;

    explicit world(context &ctx) : ctx_(ctx) {}
    ~world() {
        if (w_) fsg_world_destroy(w_);
    }
    world(const world &) = delete;
    world &operator=(const world &) = delete;

    // world::init: allocates the grid (all air) on the device.
    void init(int w, int h, uint64_t seed, const fsg_config *cfg = nullptr) {
        if (w_) fsg_world_destroy(w_);
        w_ = nullptr;
        check(fsg_world_create(ctx_.get(), w, h, seed, cfg, &w_), "fsg_world_create");
        width = w;
        height = h;
        edits_.clear();
        npx_ = (w + kPage - 1) / kPage;
        npy_ = (h + kPage - 1) / kPage;
        pages_.assign((size_t)npx_ * npy_, Page{});
        invalidate();
    }

    // Whole-grid load, e.g. from a generator or a saved world.
    void load(const std::vector<MaterialInstance> &cells) {
        invalidate();
        check(fsg_world_upload(w_, cells.data(), cells.size()), "fsg_world_upload");
    }

    MaterialInstance getTile(int x, int y) {
        if (x < 0 || x >= width || y < 0 || y >= height) {
            MaterialInstance wall{15, 0, 0};  // Tiles_TEST_SOLID
            return wall;
        }
        Page &p = page_at(x / kPage, y / kPage);
        if (p.gen != gen_) fetch(x / kPage, y / kPage, x / kPage, y / kPage);
        return p.cells[(size_t)(y % kPage) * kPage + (x % kPage)];
    }

    // Queued like the reference's direct real_tiles writes; applied before the
    // next device call, and at once to a mirrored page.
    void setTile(int x, int y, MaterialInstance type) {
        if (x < 0 || x >= width || y < 0 || y >= height) return;
        edits_.push_back(fsg_cell_edit{x, y, type});
        Page &p = page_at(x / kPage, y / kPage);
        if (p.gen == gen_) p.cells[(size_t)(y % kPage) * kPage + (x % kPage)] = type;
    }

    // Mirrors every page that meets [x0, x0+w) x [y0, y0+h), in one download.
    void prefetch(int x0, int y0, int w, int h) {
        const int xa = std::max(x0, 0), ya = std::max(y0, 0);
        const int xb = std::min(x0 + w, width), yb = std::min(y0 + h, height);
        if (xa >= xb || ya >= yb) return;
        fetch(xa / kPage, ya / kPage, (xb - 1) / kPage, (yb - 1) / kPage);
    }

    // Forgets the mirror (the grid changed on the device).
    void invalidate() { ++gen_; }

    void tick(int n = 1) {
        flush();
        invalidate();
        check(fsg_world_step(w_, n), "fsg_world_step");
    }

    void tickTemperature() {
        flush();
        invalidate();
        check(fsg_world_tick_temperature(w_), "fsg_world_tick_temperature");
    }

    // World::tickZone (world.hpp:138): only [x, x+w) x [y, y+h) ticks.
    void setTickZone(int x, int y, int w, int h) {
        flush();
        check(fsg_world_set_tick_zone(w_, x, y, w, h), "fsg_world_set_tick_zone");
    }
    void tickZone(int &x, int &y, int &w, int &h) const {
        int32_t a, b, c, d;
        check(fsg_world_get_tick_zone(w_, &a, &b, &c, &d), "fsg_world_get_tick_zone");
        x = a, y = b, w = c, h = d;
    }

    // Downloads the rectangle [x0, x0+w) x [y0, y0+h) (row-major); the pages
    // it covers whole are mirrored for getTile.
    std::vector<MaterialInstance> region(int x0, int y0, int w, int h) {
        flush();
        std::vector<MaterialInstance> out((size_t)w * h);
        check(fsg_world_download(w_, out.data(), x0, y0, w, h), "fsg_world_download");
        for (int py = (y0 + kPage - 1) / kPage; py < npy_ && py * kPage < y0 + h; py++) {
            for (int px = (x0 + kPage - 1) / kPage; px < npx_ && px * kPage < x0 + w; px++) {
                const int pw = std::min(kPage, width - px * kPage), ph = std::min(kPage, height - py * kPage);
                if (px * kPage + pw > x0 + w || py * kPage + ph > y0 + h) continue;
                Page &p = page_at(px, py);
                p.cells.resize((size_t)kPage * kPage);
                for (int r = 0; r < ph; r++)
                    std::copy_n(out.data() + (size_t)(py * kPage + r - y0) * w + (px * kPage - x0), pw,
                                p.cells.data() + (size_t)r * kPage);
                p.gen = gen_;
            }
        }
        return out;
    }

    fsg_stamp_result stampBodies(const std::vector<fsg_body> &bodies, const std::vector<MaterialInstance> &pixels) {
        flush();
        invalidate();
        fsg_stamp_result r{};
        check(fsg_world_stamp_bodies(w_, bodies.data(), (int)bodies.size(), pixels.data(), pixels.size(), &r),
              "fsg_world_stamp_bodies");
        return r;
    }

    // Pulls the pixels back into `pixels` (RigidBody::tiles).
    void unstampBodies(std::vector<MaterialInstance> &pixels) {
        flush();
        invalidate();
        check(fsg_world_unstamp_bodies(w_, pixels.data(), pixels.size()), "fsg_world_unstamp_bodies");
    }

    // The texture pass after a tick: updates the requested layers over the
    // rectangle in p and returns the movingTiles census (hadDirty, hadFire).
    fsg_render_stats updateTextures(const fsg_render_params &p) {
        flush();
        fsg_render_stats st{};
        check(fsg_world_render(w_, &p, &st), "fsg_world_render");
        return st;
    }

    // Host copy of a layer rectangle (R,G,B,A bytes per texel), the source of
    // R_UpdateImageBytes.
    std::vector<uint32_t> layer(uint32_t which, int x0, int y0, int w, int h) {
        std::vector<uint32_t> out((size_t)w * h);
        check(fsg_world_layer_download(w_, which, x0, y0, w, h, out.data()), "fsg_world_layer_download");
        return out;
    }

    // Chunk::ChunkWrite for the ncx x ncy chunks whose top-left cell is
    // (x0, y0): one .pack file image per chunk, chunk rows in order.  layer2
    // and background are region-shaped (ncx*128 wide) or null; phase holds
    // one generation phase per chunk or is null.
    std::vector<std::vector<uint8_t>> writeChunks(int x0, int y0, int ncx, int ncy, const int8_t *phase = nullptr,
                                                  const MaterialInstance *layer2 = nullptr,
                                                  const uint32_t *background = nullptr) {
        flush();
        std::vector<size_t> sizes((size_t)ncx * ncy);
        check(fsg_world_pack(w_, x0, y0, ncx, ncy, phase, layer2, background, sizes.data()), "fsg_world_pack");
        size_t total = 0;
        for (size_t v : sizes) total += v;
        std::vector<uint8_t> all(total);
        check(fsg_world_pack_download(w_, all.data(), all.size()), "fsg_world_pack_download");
        std::vector<std::vector<uint8_t>> out;
        size_t o = 0;
        for (size_t v : sizes) {
            out.emplace_back(all.begin() + o, all.begin() + o + v);
            o += v;
        }
        return out;
    }

    // Chunk::ChunkRead of one image per chunk of the region; returns each
    // chunk's status (FSG_OK, or FSG_ERR_INVALID for a corrupt image, whose
    // cells are left as they were).
    std::vector<int32_t> readChunks(int x0, int y0, int ncx, int ncy, const std::vector<std::vector<uint8_t>> &images,
                                    int8_t *phase = nullptr, MaterialInstance *layer2 = nullptr,
                                    uint32_t *background = nullptr) {
        flush();
        const size_t n = (size_t)ncx * ncy;
        if (images.size() != n) check(FSG_ERR_INVALID, "readChunks: one image per chunk");
        std::vector<size_t> sizes;
        std::vector<uint8_t> all;
        for (const auto &im : images) {
            sizes.push_back(im.size());
            all.insert(all.end(), im.begin(), im.end());
        }
        all.push_back(0);  // a valid pointer when every image is empty
        std::vector<int32_t> status(n);
        invalidate();
        check(fsg_world_unpack(w_, x0, y0, ncx, ncy, all.data(), sizes.data(), phase, layer2, background,
                               status.data()),
              "fsg_world_unpack");
        return status;
    }

    fsg_stats stats() {
        flush();
        fsg_stats s;
        check(fsg_world_stats(w_, &s), "fsg_world_stats");
        return s;
    }

    fsg_world *handle() const { return w_; }

  private:
    static constexpr int kPage = 64;
    struct Page {
        uint64_t gen = 0;  // valid while equal to gen_
        std::vector<MaterialInstance> cells;  // kPage x kPage, row-major
    };

    void flush() {
        if (edits_.empty()) return;
        check(fsg_world_set_cells(w_, edits_.data(), edits_.size()), "fsg_world_set_cells");
        edits_.clear();
    }

    Page &page_at(int px, int py) { return pages_[(size_t)py * npx_ + px]; }

    // Downloads pages [px0, px1] x [py0, py1] in one call.
    void fetch(int px0, int py0, int px1, int py1) {
        flush();
        const int x0 = px0 * kPage, y0 = py0 * kPage;
        const int w = std::min((px1 + 1) * kPage, width) - x0, h = std::min((py1 + 1) * kPage, height) - y0;
        std::vector<MaterialInstance> blk((size_t)w * h);
        check(fsg_world_download(w_, blk.data(), x0, y0, w, h), "fsg_world_download");
        for (int py = py0; py <= py1; py++) {
            for (int px = px0; px <= px1; px++) {
                Page &p = page_at(px, py);
                p.cells.resize((size_t)kPage * kPage);
                const int pw = std::min(kPage, width - px * kPage), ph = std::min(kPage, height - py * kPage);
                for (int r = 0; r < ph; r++)
                    std::copy_n(blk.data() + (size_t)(py * kPage + r - y0) * w + (px * kPage - x0), pw,
                                p.cells.data() + (size_t)r * kPage);
                p.gen = gen_;
            }
        }
    }

    context &ctx_;
    fsg_world *w_ = nullptr;
    std::vector<fsg_cell_edit> edits_;
    std::vector<Page> pages_;
    int npx_ = 0, npy_ = 0;
    uint64_t gen_ = 1;
};

}  // namespace fsg

#endif  // FSG_WORLD_HPP
