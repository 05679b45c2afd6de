/*
 * fsg_oracle.h -- CPU oracle for the deterministic falling-sand tick.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (libfsg.so) never links or calls it.
 *
 * Two CPU modes live here:
 *   det   (fsg_oracle.c)  the deterministic phase-schedule + counter-RNG tick
 *                         that the GPU must match bit-exactly;
 *   refseq (refseq.c)     a restatement of the reference's sequential,
 *                         order-dependent world::tick on a thread pool; the
 *                         timed CPU baseline and the statistical comparator.
 */
#ifndef FSG_ORACLE_H
#define FSG_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#include "../include/fsg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The benchmark material table, written independently of the product's. */
void orc_default_config(fsg_config *cfg);

uint64_t orc_splitmix64(uint64_t state);          /* next() of splitmix64 at `state` */
uint32_t orc_lowbias32(uint32_t x);
uint32_t orc_lowbias32_inv(uint32_t x);
uint64_t orc_phase_key(uint64_t seed, uint64_t g); /* key of global phase g          */
void orc_block_hash(uint64_t key, int x0, int y0, uint32_t *r, uint32_t *s);

/* One Margolus block update.  q[4] = TL, TR, BL, BR (as u64 cells);
 * returns 1 if any cell changed. */
int orc_block_rule(const fsg_config *cfg, int it, uint32_t r, uint32_t s, uint64_t q[4]);

/* One det tick (cell_iter phases + temperature if the cadence says so),
 * in place on a W*H row-major grid of u64 cells. */
int orc_tick(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed, uint64_t tick);
int orc_run(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed,
            uint64_t tick0, int nticks);
/* The same, limited to a tick zone zone[0..3] = x, y, w, h (World::tickZone):
 * a block acts iff one of its cells lies in the zone, and only zone cells
 * take the temperature step.  NULL = the whole world. */
int orc_tick_zone(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed, uint64_t tick,
                  const int32_t *zone);
int orc_run_zone(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed, uint64_t tick0,
                 int nticks, const int32_t *zone);
/* One temperature step (5-point, 1/8, floor; out-of-world neighbour = self). */
int orc_temperature(uint64_t *cells, int W, int H, const fsg_config *cfg);
int orc_temperature_zone(uint64_t *cells, int W, int H, const fsg_config *cfg, const int32_t *zone);

/* Dirty-rect predicate: 1 if any cell of the rectangle [x0,x1)x[y0,y1) or
 * any pair of cells within it could change under some RNG draw, or any
 * temperature stencil centred in [cx0,cx1)x[cy0,cy1) is non-zero. */
int orc_region_unstable(const uint64_t *cells, int W, int H, const fsg_config *cfg,
                        int x0, int y0, int x1, int y1,
                        int cx0, int cy0, int cx1, int cy1);

/* Rigid-body stamp / unstamp in deterministic claim rounds. */
int orc_stamp(uint64_t *cells, int W, int H, const fsg_config *cfg, const fsg_body *bodies,
              int nbodies, const uint64_t *pixels, size_t npixels, uint8_t *placed,
              uint32_t *n_placed, uint32_t *n_displaced,
              fsg_cell_edit *displaced_out, uint32_t displaced_cap,
              uint32_t *hits_sand, uint32_t *hits_soup);
int orc_unstamp(uint64_t *cells, int W, int H, const fsg_config *cfg, const fsg_body *bodies,
                int nbodies, uint64_t *pixels, size_t npixels);
int orc_body_pixel_xy(const fsg_body *b, int tx, int ty, int *wx, int *wy);

/* world::physicsCheck (world.cpp:3330-3429) for seeds in order (each acts on
 * the grid the previous ones left).  Returns -1 without changing anything if
 * the body images need more than pixels_cap cells (*pixels_needed). */
int orc_physics_check(uint64_t *cells, int W, int H, const fsg_config *cfg, fsg_island *seeds, int n,
                      uint64_t *pixels_out, size_t pixels_cap, size_t *pixels_needed);

/* Texture pass (game.cpp:1994-2067) over p's rectangle: the cells, emission
 * and fire layers (any may be NULL), and with prev (the state before the
 * last tick) and st the movingTiles census.  Layers are updated in place. */
int orc_render(const uint64_t *cells, const uint64_t *prev, int W, int H, const fsg_config *cfg,
               const fsg_render_params *p, uint32_t *lc, uint32_t *le, uint32_t *lf,
               fsg_render_stats *st);

/* Chunk .pack images (Chunk::ChunkWrite chunk.cpp:219-330, Chunk::ChunkRead
 * chunk.cpp:74-217).  A chunk is 128x128 cells (const.h:19-20). */
#define ORC_PACK_NA 393216 /* 2 * 128 * 128 MaterialInstanceData records of 12 bytes */
#define ORC_PACK_NB 65536  /* 128 * 128 u32 background */
#define ORC_PACK_HDR 17
#define ORC_LZ4_BOUND(n) ((n) + (n) / 255 + 16) /* LZ4_compressBound (lz4.h) */

/* The records of one chunk (chunk.cpp:238-242): raw_a = {u32 index, u32
 * colour, i16 temperature, pad byte 10 = MOVED, pad byte 11 = 0} of tiles
 * then layer2, raw_b = the background.  index = map->to_file[material id]
 * (map NULL: the id itself); FSG_PACK_NO_MOVED writes pad byte 10 as 0.
 * Inputs are chunk windows with row strides in elements; layer2 / bg may be
 * NULL (zero records). */
void orc_chunk_records(const uint64_t *tiles, size_t tstride, const uint64_t *layer2, size_t lstride,
                       const uint32_t *bg, size_t bstride, uint8_t *raw_a, uint8_t *raw_b,
                       const fsg_pack_map *map);
/* LZ4 block decoding with the acceptance rules of lz4.c's
 * LZ4_decompress_safe for a full block (lz4.c:1949-2360): decoded size, or -1. */
int orc_lz4_decode(const uint8_t *src, int n, uint8_t *dst, int cap);
/* The parse of the library's encoder for an n-byte block of rec-byte
 * records, 128 records a row: each byte is a match candidate against the
 * byte one record back, else one row back; runs of 4+ such bytes inside a
 * row become matches, the rest literals; the last 12 bytes are literals.
 * Returns the compressed size, or -1 if cap is too small. */
int orc_lz4_encode_rows(const uint8_t *raw, int n, int rec, uint8_t *dst, int cap);
/* One .pack image of the records (header + block A + block B); returns its
 * size.  img holds at least 17 + ORC_LZ4_BOUND(NA) + ORC_LZ4_BOUND(NB). */
size_t orc_pack_chunk(const uint8_t *raw_a, const uint8_t *raw_b, int8_t phase, uint8_t *img);
/* ChunkRead of one image into chunk windows (layer2 / bg may be NULL):
 * FSG_OK, or FSG_ERR_INVALID for bad header sizes, a block that does not
 * decode to its full size or a record index that names no material (index
 * >= 256 or map->from_file[index] < 0; map NULL: index >= n_materials)
 * (nothing written then; *phase is still set once the header has been
 * read).  MOVED comes from bit 0 of pad byte 10 unless FSG_PACK_NO_MOVED. */
int orc_unpack_chunk(const uint8_t *img, size_t n, uint32_t n_materials, uint64_t *tiles, size_t tstride,
                     uint64_t *layer2, size_t lstride, uint32_t *bg, size_t bstride, int8_t *phase,
                     const fsg_pack_map *map);

/* The reference's own pixel-major stamp and unstamp loops (game.cpp:1768-1812,
 * 1908-1965): each pixel, tx outer and ty inner, takes its first free
 * checkDirs offset before the next pixel is tried.  Test pins of the
 * claim-round forms above, not the build's definition. */
int orc_refstamp(uint64_t *cells, int W, int H, const fsg_config *cfg, const fsg_body *bodies, int nbodies,
                 const uint64_t *pixels, size_t npixels, uint8_t *placed, uint32_t *n_placed,
                 uint32_t *n_displaced, fsg_cell_edit *displaced_out, uint32_t displaced_cap);
int orc_refunstamp(uint64_t *cells, int W, int H, const fsg_config *cfg, const fsg_body *bodies, int nbodies,
                   uint64_t *pixels, size_t npixels);

/* Marching-squares outline of a binary mask (fsg_mesh_oracle.c): the
 * vertex value of MarchingSquares::value, FindPerimeter's runs, simplify,
 * and updateChunkMesh's loop scan over a whole mask. */
int orc_ms_value(const uint8_t *m, int W, int H, int x, int y);
int orc_find_perimeter(const uint8_t *m, int W, int H, int x0, int y0, int32_t *dirs, int cap);
size_t orc_simplify(const float *pts, size_t n, float tol, float *out);
int orc_mesh_mask(const uint8_t *m, int W, int H, fsg_mesh_loop *loops, uint32_t loop_cap, float *verts,
                  uint32_t vert_cap, uint32_t *nloops, uint32_t *nverts);

/* world::tickCells (fsg_particle_oracle.c): one pass over the particle list
 * in order, landing into cells; the list is compacted in place. */
int orc_spiral(int32_t *offs);
int orc_particles_tick(uint64_t *cells, int W, int H, const fsg_config *cfg, fsg_particle *p, size_t n,
                       size_t *alive, size_t *landed);
/* The same with a tick zone (NULL = the whole world): a particle whose cell
 * lies outside it keeps its place (world.cpp:2061). */
int orc_particles_tick_zone(uint64_t *cells, int W, int H, const fsg_config *cfg, fsg_particle *p, size_t n,
                            size_t *alive, size_t *landed, const int32_t *zone);

void orc_counts(const uint64_t *cells, size_t n, uint64_t counts[16]);
uint64_t orc_checksum(const uint64_t *cells, size_t n);

/* refseq: the reference's sequential chunk-checkerboard tick. */
int refseq_run(uint64_t *cells, int W, int H, const fsg_config *cfg, uint64_t seed,
               int nticks, int threads);

#ifdef __cplusplus
}
#endif
#endif
