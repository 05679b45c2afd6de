// Drives the C++ facade (include/fsg_world.hpp) the way game::tick drives
// ME::world: load a world, poke tiles with setTile, tick, read with getTile.
// Prints one JSON line with the checksum and material counts.
#include <cstdio>
#include <vector>

#include "fsg_world.hpp"

int main() {
    try {
        fsg::context ctx(0);
        fsg::world world(ctx);
        const int W = 256, H = 192;
        world.init(W, H, 7);
        std::vector<fsg::MaterialInstance> cells((size_t)W * H);
        fsg::check(fsg_generate(2, W, H, 7, cells.data()), "fsg_generate");
        world.load(cells);
        fsg::MaterialInstance sand{2, 320, 0xFF50C8E6u};
        for (int x = 100; x < 110; x++) world.setTile(x, 5, sand);
        world.setTile(-3, 5, sand);  // ignored, like world::setTile
        world.tick(10);
        const fsg::MaterialInstance wall = world.getTile(-1, 0);
        const fsg::MaterialInstance t = world.getTile(50, 50);
        const fsg_stats s = world.stats();
        // the texture pass: cells layer of the whole world, then one texel read back
        fsg_render_params rp{};
        rp.layers = FSG_LAYER_CELLS;
        rp.fire_mat = 255;
        world.updateTextures(rp);
        const std::vector<uint32_t> texel = world.layer(FSG_LAYER_CELLS, 50, 50, 1, 1);
        // chunk files: write the two 128x128 chunks of the top rows, read them back over other cells
        const fsg_stats before_files = world.stats();
        const std::vector<std::vector<uint8_t>> files = world.writeChunks(0, 0, 2, 1);
        for (int x = 0; x < 128; x++) world.setTile(x, 3, sand);
        const std::vector<int32_t> st = world.readChunks(0, 0, 2, 1, files);
        const bool files_ok = st[0] == FSG_OK && st[1] == FSG_OK && world.stats().checksum == before_files.checksum;
        // getTile's mirror: cell by cell equal to region(), setTile seen at
        // once, and a tick drops it (reads after the tick see the new grid)
        auto same = [](const fsg::MaterialInstance &a, const fsg::MaterialInstance &b) {
            return a.mat == b.mat && a.temp == b.temp && a.rgba == b.rgba;
        };
        bool mirror_ok = true;
        for (int pass = 0; pass < 2; pass++) {
            const std::vector<fsg::MaterialInstance> reg = world.region(0, 0, W, H);
            for (int y = 0; y < H; y += (pass ? 1 : 3))
                for (int x = 0; x < W; x += (pass ? 3 : 1))
                    mirror_ok = mirror_ok && same(world.getTile(x, y), reg[(size_t)y * W + x]);
            world.setTile(131, 67, sand);
            mirror_ok = mirror_ok && same(world.getTile(131, 67), sand);
            world.tick(5);
        }
        world.prefetch(60, 60, 100, 50);
        const std::vector<fsg::MaterialInstance> reg3 = world.region(60, 60, 100, 50);
        for (int y = 0; y < 50; y++)
            for (int x = 0; x < 100; x++) mirror_ok = mirror_ok && same(world.getTile(60 + x, 60 + y), reg3[(size_t)y * 100 + x]);
        std::printf("{\"checksum\": %llu, \"files_ok\": %d, \"mirror_ok\": %d, \"wall\": %u, \"tile_50_50\": %llu, "
                    "\"texel_50_50\": %u, \"counts\": [",
                    (unsigned long long)s.checksum, files_ok ? 1 : 0, mirror_ok ? 1 : 0, (unsigned)wall.mat,
                    (unsigned long long)(t.mat | ((uint64_t)(uint16_t)t.temp << 16) | ((uint64_t)t.rgba << 32)),
                    (unsigned)texel[0]);
        for (int i = 0; i < 8; i++) std::printf("%llu%s", (unsigned long long)s.count[i], i < 7 ? ", " : "");
        std::printf("]}\n");
        return 0;
    } catch (const fsg::error &e) {
        std::fprintf(stderr, "fsg error: %s\n", e.what());
        return 2;
    }
}
