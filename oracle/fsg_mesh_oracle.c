/* fsg_mesh_oracle.c -- CPU restatement of the reference's outline step of
 * world::updateChunkMesh (world.cpp:722-895): the perimeter scan over a
 * binary mask, MarchingSquares::value / FindPerimeter (physics_math.cpp:1866-1965),
 * the edgeSeen marking, the mesh vertices, simplify (Ramer-Douglas-Peucker,
 * physics_math.cpp:1766-1843) and TPPLPoly::GetOrientation
 * (physics_math.cpp:221-232).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the device mesher, never linked
 * into the product. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fsg_oracle.h"

/* isSet (physics_math.cpp:1866): vertex coordinates are one more than cell
 * coordinates; outside the mask is unset. */
static int is_set(const uint8_t *m, int W, int H, int x, int y) {
    return (x <= 0 || x > W || y <= 0 || y > H) ? 0 : m[(size_t)(y - 1) * W + (x - 1)] != 0;
}

int orc_ms_value(const uint8_t *m, int W, int H, int x, int y) {
    return is_set(m, W, H, x, y) | is_set(m, W, H, x + 1, y) << 1 | is_set(m, W, H, x, y + 1) << 2 |
           is_set(m, W, H, x + 1, y + 1) << 3;
}

/* FindPerimeter (physics_math.cpp:1881-1965): directions as (dx, dy) runs,
 * equal consecutive directions merged; y grows up in direction space and
 * down in vertex space.  Returns the number of runs, 0 if the start does
 * not lie on a perimeter, -1 if dirs is too small or the walk does not
 * close within 4 (W+1) (H+1) steps. */
int orc_find_perimeter(const uint8_t *m, int W, int H, int x0, int y0, int32_t *dirs, int cap) {
    if (x0 < 0) x0 = 0;
    if (x0 > W) x0 = W;
    if (y0 < 0) y0 = 0;
    if (y0 > H) y0 = H;
    const int v0 = orc_ms_value(m, W, H, x0, y0);
    if (v0 == 0 || v0 == 15) return 0;
    int x = x0, y = y0, px = 0, py = 0, n = 0;
    const long limit = 4L * (W + 1) * (H + 1);
    for (long step = 0;; step++) {
        if (step > limit) return -1;
        int dx = 0, dy = 0;
        switch (orc_ms_value(m, W, H, x, y)) {
        case 1: case 5: case 13: dy = 1; break;           /* North */
        case 2: case 3: case 7: dx = 1; break;            /* East */
        case 4: case 12: case 14: dx = -1; break;         /* West */
        case 8: case 10: case 11: dy = -1; break;         /* South */
        case 6:                                           /* previous North ? West : East */
            if (px == 0 && py == 1) dx = -1;
            else dx = 1;
            break;
        case 9:                                           /* previous East ? North : South */
            if (px == 1 && py == 0) dy = 1;
            else dy = -1;
            break;
        default: return -1;                               /* "Illegal state" */
        }
        if (n > 0 && dx == px && dy == py) {
            dirs[2 * (n - 1)] += dx;
            dirs[2 * (n - 1) + 1] += dy;
        } else {
            if (n >= cap) return -1;
            dirs[2 * n] = dx;
            dirs[2 * n + 1] = dy;
            n++;
            px = dx;
            py = dy;
        }
        x += dx;
        y -= dy;
        if (x == x0 && y == y0) break;
    }
    return n;
}

/* pDistance (physics_math.cpp:1814-1843) in f32, no fused operations. */
static float p_distance(float x, float y, float x1, float y1, float x2, float y2) {
    volatile float A = x - x1, B = y - y1, C = x2 - x1, D = y2 - y1;
    volatile float ac = A * C, bd = B * D, cc = C * C, dd = D * D;
    volatile float dot = ac + bd, len_sq = cc + dd;
    float param = -1;
    if (len_sq != 0) param = dot / len_sq;
    float xx, yy;
    if (param < 0) {
        xx = x1;
        yy = y1;
    } else if (param > 1) {
        xx = x2;
        yy = y2;
    } else {
        volatile float pc = param * C, pd = param * D;
        xx = x1 + pc;
        yy = y1 + pd;
    }
    volatile float dx = x - xx, dy = y - yy;
    volatile float dx2 = dx * dx, dy2 = dy * dy;
    volatile float s = dx2 + dy2;
    return sqrtf(s);
}

/* simplify_section (physics_math.cpp:1766-1797): the recursion made an
 * explicit stack.  The reference's "keep 2 points" guard compares the full
 * point count (its omitted counter only grows in branches that do not
 * recurse). */
static void simplify_section(const float *pts, size_t n, float tol, size_t i0, size_t j0, uint8_t *keep,
                             size_t *stack) {
    if (n <= 2) return;
    size_t sp = 0;
    stack[sp++] = i0;
    stack[sp++] = j0;
    while (sp) {
        const size_t j = stack[--sp], i = stack[--sp];
        if (i + 1 == j) continue;
        float maxd = -1.0f;
        size_t maxk = i;
        for (size_t k = i + 1; k < j; k++) {
            const float d = p_distance(pts[2 * k], pts[2 * k + 1], pts[2 * i], pts[2 * i + 1], pts[2 * j], pts[2 * j + 1]);
            if (d > maxd) {
                maxd = d;
                maxk = k;
            }
        }
        if (maxd <= tol) {
            for (size_t k = i + 1; k < j; k++) keep[k] = 0;
        } else {
            /* the reference recurses into (i, maxk) first; the two halves
             * touch disjoint marks, so the order does not matter */
            stack[sp++] = maxk;
            stack[sp++] = j;
            stack[sp++] = i;
            stack[sp++] = maxk;
        }
    }
}

size_t orc_simplify(const float *pts, size_t n, float tol, float *out) {
    uint8_t *keep = (uint8_t *)malloc(n ? n : 1);
    size_t *stack = (size_t *)malloc((4 * n + 8) * sizeof(size_t));
    memset(keep, 1, n);
    if (n) simplify_section(pts, n, tol, 0, n - 1, keep, stack);
    size_t m = 0;
    for (size_t k = 0; k < n; k++)
        if (keep[k]) {
            out[2 * m] = pts[2 * k];
            out[2 * m + 1] = pts[2 * k + 1];
            m++;
        }
    free(keep);
    free(stack);
    return m;
}

/* The outline step of updateChunkMesh over one mask (world.cpp:738-895),
 * up to the polygons handed to TPPLPartition: every kept loop's vertices in
 * the order of worldMesh (the reference fills TPPLPoly in reverse) and
 * whether GetOrientation calls it clockwise (a hole: SetHole(true)).
 * Returns FSG_OK, or FSG_ERR_INVALID when an output array is too small. */
int orc_mesh_mask(const uint8_t *m, int W, int H, fsg_mesh_loop *loops, uint32_t loop_cap, float *verts,
                  uint32_t vert_cap, uint32_t *nloops, uint32_t *nverts) {
    const size_t size = (size_t)W * H;
    uint8_t *seen = (uint8_t *)calloc(size ? size : 1, 1);
    const int cap = (int)(4 * size + 8);
    int32_t *dirs = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)cap);
    float *mesh = (float *)malloc(sizeof(float) * 2 * (size_t)cap);
    float *simp = (float *)malloc(sizeof(float) * 2 * (size_t)cap);
    uint32_t nl = 0, nv = 0;
    int rc = FSG_OK;
    size_t look = 0;
    for (;;) {
        /* the next solid cell from lookIndex whose right, lower and lower
         * right neighbours are not all solid (world.cpp:768-788) */
        long edge = -1;
        for (size_t i = look; i < size; i++) {
            if (!m[i]) continue;
            const int x = (int)(i % W), y = (int)(i / W);
            int nb = 0;
            if (x + 1 < W) nb += m[i + 1] != 0;
            if (y + 1 < H) nb += m[i + W] != 0;
            if (y + 1 < H && x + 1 < W) nb += m[i + W + 1] != 0;
            if (nb != 3) {
                edge = (long)i;
                break;
            }
        }
        if (edge < 0) break;
        const int lx = (int)(edge % W), ly = (int)(edge / W);
        look = (size_t)edge + 1;
        if (seen[edge]) continue;
        const int val = orc_ms_value(m, W, H, lx, ly);
        if (val == 0 || val == 15) continue;
        const int nr = orc_find_perimeter(m, W, H, lx, ly, dirs, cap);
        if (nr < 0) {
            rc = FSG_ERR_INVALID;
            break;
        }
        /* edgeSeen along each run from its start vertex, then the run's end
         * vertex into the mesh (world.cpp:843-868) */
        float lastx = (float)lx, lasty = (float)ly;
        for (int r = 0; r < nr; r++) {
            const int dx = dirs[2 * r], dy = dirs[2 * r + 1];
            const int nx = abs(dx) > 1 ? abs(dx) : 1, ny = abs(dy) > 1 ? abs(dy) : 1;
            for (int ix = 0; ix < nx; ix++)
                for (int iy = 0; iy < ny; iy++) {
                    int ilx = (int)(lastx + (float)(ix * (dx < 0 ? -1 : 1)));
                    int ily = (int)(lasty - (float)(iy * (dy < 0 ? -1 : 1)));
                    if (ilx < 0) ilx = 0;
                    if (ilx >= W) ilx = W - 1;
                    if (ily < 0) ily = 0;
                    if (ily >= H) ily = H - 1;
                    seen[(size_t)ilx + (size_t)ily * W] = 1;
                }
            lastx += (float)dx;
            lasty -= (float)dy;
            mesh[2 * r] = lastx;
            mesh[2 * r + 1] = lasty;
        }
        const size_t ns = orc_simplify(mesh, (size_t)nr, 1.0f, simp);
        if (ns < 3) continue; /* "1x1 pixel that breaks everything" */
        if (nl >= loop_cap || nv + ns > vert_cap) {
            rc = FSG_ERR_INVALID;
            break;
        }
        /* GetOrientation of the reversed point list: the shoelace sum in
         * integers (every vertex is a whole number) */
        int64_t area = 0;
        for (size_t a = 0; a < ns; a++) {
            const size_t p = ns - 1 - a, q = (a + 1 == ns) ? ns - 1 : ns - 2 - a;
            area += (int64_t)simp[2 * p] * (int64_t)simp[2 * q + 1] - (int64_t)simp[2 * p + 1] * (int64_t)simp[2 * q];
        }
        loops[nl].first = nv;
        loops[nl].count = (uint32_t)ns;
        loops[nl].hole = area < 0;
        loops[nl].start_x = lx;
        loops[nl].start_y = ly;
        loops[nl].region = 0;
        memcpy(verts + 2 * (size_t)nv, simp, sizeof(float) * 2 * ns);
        nv += (uint32_t)ns;
        nl++;
    }
    free(seen);
    free(dirs);
    free(mesh);
    free(simp);
    *nloops = nl;
    *nverts = nv;
    return rc;
}
