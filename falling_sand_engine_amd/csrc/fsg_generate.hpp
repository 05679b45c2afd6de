// fsg_generate.hpp -- the benchmark world generator's plan, shared by the
// host generator (fsg_generate, fsg_host.cpp) and the device generator
// (fsg_world_generate, fsg_generate.hip) so both draw the same discs.
#pragma once
#include <stdint.h>

#include <vector>

struct fsg_gen_disc {
    int32_t cx, cy, r, m;  // centre, radius, material
};

// Disc list of a basin kind (1, 2, 4) in drawing order (empty for kind 3),
// the base colours of the 8 materials and the temperatures make() writes.
int fsg_gen_plan(int kind, int W, int H, uint64_t seed, std::vector<fsg_gen_disc> &discs, uint32_t rgba[8],
                 int16_t &ambient, int16_t &lava_temp);
