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
#ifndef FSG_WORLD_HPP
#define FSG_WORLD_HPP

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
    }

    // Whole-grid load, e.g. from a generator or a saved world.
    void load(const std::vector<MaterialInstance> &cells) {
        check(fsg_world_upload(w_, cells.data(), cells.size()), "fsg_world_upload");
    }

    MaterialInstance getTile(int x, int y) {
        if (x < 0 || x >= width || y < 0 || y >= height) {
            MaterialInstance wall{15, 0, 0};  // Tiles_TEST_SOLID
            return wall;
        }
        flush();
        MaterialInstance c;
        check(fsg_world_download(w_, &c, x, y, 1, 1), "fsg_world_download");
        return c;
    }

    // Queued like the reference's direct real_tiles writes; applied before the next tick.
    void setTile(int x, int y, MaterialInstance type) {
        if (x < 0 || x >= width || y < 0 || y >= height) return;
        edits_.push_back(fsg_cell_edit{x, y, type});
    }

    void tick(int n = 1) {
        flush();
        check(fsg_world_step(w_, n), "fsg_world_step");
    }

    void tickTemperature() {
        flush();
        check(fsg_world_tick_temperature(w_), "fsg_world_tick_temperature");
    }

    // Downloads the rectangle [x0, x0+w) x [y0, y0+h) (row-major).
    std::vector<MaterialInstance> region(int x0, int y0, int w, int h) {
        flush();
        std::vector<MaterialInstance> out((size_t)w * h);
        check(fsg_world_download(w_, out.data(), x0, y0, w, h), "fsg_world_download");
        return out;
    }

    fsg_stamp_result stampBodies(const std::vector<fsg_body> &bodies, const std::vector<MaterialInstance> &pixels) {
        flush();
        fsg_stamp_result r{};
        check(fsg_world_stamp_bodies(w_, bodies.data(), (int)bodies.size(), pixels.data(), pixels.size(), &r),
              "fsg_world_stamp_bodies");
        return r;
    }

    // Pulls the pixels back into `pixels` (RigidBody::tiles).
    void unstampBodies(std::vector<MaterialInstance> &pixels) {
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
    void flush() {
        if (edits_.empty()) return;
        check(fsg_world_set_cells(w_, edits_.data(), edits_.size()), "fsg_world_set_cells");
        edits_.clear();
    }

    context &ctx_;
    fsg_world *w_ = nullptr;
    std::vector<fsg_cell_edit> edits_;
};

}  // namespace fsg

#endif  // FSG_WORLD_HPP
