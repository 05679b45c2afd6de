/* fsg_particle_oracle.c -- CPU restatement of world::tickCells
 * (world.cpp:2030-2189) over a list of CellData (game_utils/cells.h:15-36):
 * the particles in list order, each seeing the cells the particles before it
 * landed in during the same pass.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the device particle pass, never
 * linked into the product. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fsg_oracle.h"

static int cell_type(const uint64_t *cells, const fsg_config *cfg, size_t i) {
    const uint32_t m = (uint32_t)(cells[i] & 15u);
    return m < cfg->n_materials ? cfg->materials[m].type : FSG_SOLID;
}

static uint64_t cell_of(const fsg_cell *c) {
    return (uint64_t)c->mat | ((uint64_t)(uint16_t)c->temp << 16) | ((uint64_t)c->rgba << 32);
}

/* The square spiral of tickCells' landing search (world.cpp:2113-2139):
 * X = Y = 32, 1024 steps from (0, 0) starting downwards (dy = -1), turning
 * where x == y, or x < 0 and x == -y, or x > 0 and x == 1 - y; offsets with
 * |x|, |y| <= 16 are tried.  Writes the tried offsets in order, returns
 * their count. */
int orc_spiral(int32_t *offs) {
    int x = 0, y = 0, dx = 0, dy = -1, n = 0;
    for (int j = 0; j < 32 * 32; j++) {
        if (-16 <= x && x <= 16 && -16 <= y && y <= 16) {
            offs[2 * n] = x;
            offs[2 * n + 1] = y;
            n++;
        }
        if (x == y || (x < 0 && x == -y) || (x > 0 && x == 1 - y)) {
            const int t = dx;
            dx = -dy;
            dy = t;
        }
        x += dx;
        y += dy;
    }
    return n;
}

int orc_particles_tick_zone(uint64_t *cells, int W, int H, const fsg_config *cfg, fsg_particle *p, size_t n,
                            size_t *alive, size_t *landed, const int32_t *zone) {
    if (zone && (zone[0] < 0 || zone[1] < 0 || zone[2] < 1 || zone[3] < 1 || zone[2] > W - zone[0] ||
                 zone[3] > H - zone[1]))
        return FSG_ERR_INVALID;
    static int32_t spiral[2 * 1024];
    static int nspiral = -1;
    if (nspiral < 0) nspiral = orc_spiral(spiral);
    size_t out = 0, nl = 0;
    for (size_t i = 0; i < n; i++) {
        fsg_particle c = p[i];
        int keep = 1;
        if ((c.flags & FSG_PARTICLE_TEMPORARY) && c.lifetime <= 0) continue;
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
        if (c.x < 0 || (int)c.x >= W || c.y < 0 || (int)c.y >= H) continue;
        /* outside the tick zone: kept as it is, the target pull included */
        if (zone && !(lx >= zone[0] && ly >= zone[1] && lx < zone[0] + zone[2] && ly < zone[1] + zone[3])) {
            p[out++] = c;
            continue;
        }
        c.vx += c.ax;
        c.vy += c.ay;
        const int div = (int)((fabsf(c.vx) + fabsf(c.vy)) + 1);
        const float dvx = c.vx / (float)div, dvy = c.vy / (float)div;
        int done = 0;
        for (int k = 0; k < div && !done; k++) {
            c.x += dvx;
            c.y += dvy;
            if (c.x < 0 || (int)c.x >= W || c.y < 0 || (int)c.y >= H) {
                keep = 0;
                done = 1;
                break;
            }
            const size_t at = (size_t)(int)c.x + (size_t)(int)c.y * (size_t)W;
            if ((c.flags & FSG_PARTICLE_PHASE) || cell_type(cells, cfg, at) == FSG_AIR) continue;
            /* inObjectState (world.cpp:2081-2092): no OBJECT cells on this grid */
            if (c.in_object_state == 0 || c.in_object_state == 1) c.in_object_state = 2;
            done = 1;
            if (c.flags & FSG_PARTICLE_TEMPORARY) {
                keep = 0;
                break;
            }
            const size_t home = (size_t)lx + (size_t)ly * (size_t)W;
            if (cell_type(cells, cfg, home) != FSG_AIR) {
                int ok = 0;
                for (int s = 0; s < nspiral; s++) {
                    const int cx = (int)(c.x + (float)spiral[2 * s]), cy = (int)(c.y + (float)spiral[2 * s + 1]);
                    /* the reference indexes real_tiles here without a bounds
                     * check (world.cpp:2126); offsets outside the world are
                     * skipped instead of read out of range */
                    if (cx < 0 || cy < 0 || cx >= W || cy >= H) continue;
                    const size_t ci = (size_t)cx + (size_t)cy * (size_t)W;
                    if (cell_type(cells, cfg, ci) == FSG_AIR) {
                        cells[ci] = cell_of(&c.tile);
                        ok = 1;
                        break;
                    }
                }
                if (ok) {
                    keep = 0;
                    nl++;
                } else {
                    c.vy = -4.0f;
                    c.y -= 16.0f;
                }
            } else {
                cells[home] = cell_of(&c.tile);
                keep = 0;
                nl++;
            }
        }
        if (!keep) continue;
        if (!done && c.lifetime > 0) c.lifetime--;
        p[out++] = c;
    }
    /* erase_if(cells, y > height) (world.cpp:2190) */
    size_t m = 0;
    for (size_t i = 0; i < out; i++)
        if (!(p[i].y > (float)H)) p[m++] = p[i];
    if (alive) *alive = m;
    if (landed) *landed = nl;
    return FSG_OK;
}

int orc_particles_tick(uint64_t *cells, int W, int H, const fsg_config *cfg, fsg_particle *p, size_t n,
                       size_t *alive, size_t *landed) {
    return orc_particles_tick_zone(cells, W, H, cfg, p, n, alive, landed, NULL);
}
