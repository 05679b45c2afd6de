/*
 * fsg.h -- C-ABI of the MI355X-native falling-sand tick.
 *
 * This is the drop-in boundary for the per-cell simulation path of
 * cstom4994/falling_sand_engine.  The reference has no FFI on this path: its
 * boundary is the C++ class ME::world (source/engine/world.hpp:63-193) plus
 * the rigid-body stamp/unstamp loops in game::tick
 * (source/engine/game.cpp:1760-1815, 1896-1983) that index world::real_tiles
 * directly.  Every entry point below names the reference member or loop it
 * replaces.  A C++ facade with the reference's method names (init, tick,
 * tickTemperature, getTile, setTile) sits on top of this ABI in
 * include/fsg_world.hpp.
 *
 * Conventions (all entry points):
 *   - extern "C", plain pointers and sizes, no C++ or torch types;
 *   - every function returns an int status: FSG_OK (0) or a negative
 *     fsg_status; nothing throws across the ABI;
 *   - handles are opaque; the library owns all device memory; host buffers
 *     are caller-owned and copied;
 *   - one HIP stream per context; calls on one world are serialised by the
 *     caller; different worlds may be driven from different threads (the
 *     serving queue has its own mutex).
 */
#ifndef FSG_H
#define FSG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSG_ABI_VERSION 1

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
typedef enum fsg_status {
    FSG_OK = 0,
    FSG_ERR_INVALID = -1,   /* bad argument (null handle, size, id, ...)    */
    FSG_ERR_NOMEM = -2,     /* device or host allocation failed             */
    FSG_ERR_DEVICE = -3,    /* a HIP runtime call failed / no device        */
    FSG_ERR_BUDGET = -4,    /* admission refused: HBM state budget exceeded */
    FSG_ERR_STATE = -5,     /* call not valid in the object's current state */
    FSG_ERR_CHECKSUM = -6,  /* resident state checksum mismatch             */
    FSG_ERR_TIMEOUT = -7,   /* request deadline passed                      */
    FSG_ERR_CANCELLED = -8, /* request cancelled by the caller              */
    FSG_ERR_NOTFOUND = -9,  /* unknown request / world id                   */
    FSG_ERR_AGAIN = -10     /* nothing ready yet (poll)                     */
} fsg_status;

const char *fsg_status_string(int status);

/* ------------------------------------------------------------------------ */
/* Cell: 8 bytes, replaces MaterialInstance (game_datastruct.hpp:207-225).   */
/*   mat  bits 0-7  material id (Material::id)                               */
/*        bit  8    MOVED flag (MaterialInstance::moved, sand friction)      */
/*        bits 9-15 zero                                                     */
/*   temp temperature in 1/16 degree (MaterialInstance::temperature, i16)    */
/*   rgba colour, byte order R,G,B,A in memory (MaterialInstance::color +    */
/*        Material::alpha)                                                   */
/* Grid is row-major, index x + y*width, y grows downward (world.cpp:1001).  */
/* ------------------------------------------------------------------------ */
typedef struct fsg_cell {
    uint16_t mat;
    int16_t temp;
    uint32_t rgba;
} fsg_cell;

#define FSG_MAT_ID_MASK 0x00FFu
#define FSG_MOVED 0x0100u
#define FSG_MAX_MATERIALS 15 /* ids 0..14; id 15 is reserved for the world wall */
#define FSG_MAX_SIDE 32768   /* worlds are at most 32768 cells on a side       */

/* PhysicsType (game_datastruct.hpp:87-95); PASSABLE/OBJECT are not used. */
typedef enum fsg_phys {
    FSG_AIR = 0,
    FSG_SOLID = 1,
    FSG_SAND = 2,
    FSG_SOUP = 3,
    FSG_GAS = 4
} fsg_phys;

/* Material (game_datastruct.hpp:130-168).  density is replaced by an integer
 * rank (0 lightest); a heavier SAND/SOUP cell sinks into any non-SOLID cell
 * of lower rank.  iterations gates the cell_iter loop exactly as
 * world.cpp:1093-1096 does (a cell acts in iteration it iff it < iterations). */
typedef struct fsg_material {
    char name[16];
    uint8_t type;        /* fsg_phys                                       */
    uint8_t rank;        /* density rank 0..14                             */
    uint8_t iterations;  /* Material::iterations                           */
    uint8_t slipperyness;/* Material::slipperyness (documentation only)    */
    int16_t add_temp;    /* Material::addTemp, 1/16 degree per temp step   */
    uint16_t reserved;
    uint32_t rgba;       /* base colour of cells created by reactions      */
} fsg_material;

/* Simulation config: the GlobalDEF fields on this path (cvar.hpp:38-87,
 * data/scripts/global.lua:34-47) plus the build-defined reaction roles. */
typedef struct fsg_config {
    uint32_t cell_iter;      /* GlobalDEF::cell_iter (3)                    */
    uint32_t reactions;      /* 0/1: water+lava, wood burn                  */
    uint32_t temp_every;     /* 0 = temperature off; run when               */
    uint32_t temp_phase;     /*   tick % temp_every == temp_phase           */
    uint32_t chunk_skip;     /* 0/1: exact dirty-rect skipping, 64x64 chunks */
    int32_t wood_burn_temp;  /* wood turns to steam above this (1/16 deg)   */
    uint8_t mat_air, mat_water, mat_lava, mat_steam, mat_stone, mat_wood;
    uint8_t pad[2];
    uint32_t n_materials;    /* entries used in materials[]                 */
    fsg_material materials[FSG_MAX_MATERIALS];
} fsg_config;

/* Fills *cfg with the 8-material table of the benchmark worlds
 * (air, stone, sand, water, oil, steam, lava, wood), cell_iter 3,
 * reactions on, temperature every tick, chunk skipping off. */
int fsg_default_config(fsg_config *cfg);

/* ------------------------------------------------------------------------ */
/* Context: one device + one stream + an HBM state budget.                   */
/* Replaces WorldSystem's thread pools (world.hpp:56-61, world.cpp:59-61).   */
/* ------------------------------------------------------------------------ */
typedef struct fsg_ctx fsg_ctx;
typedef struct fsg_world fsg_world;

typedef struct fsg_ctx_config {
    int device;               /* HIP device ordinal                          */
    uint64_t hbm_budget;      /* bytes of world state this ctx may hold; 0 = unlimited */
} fsg_ctx_config;

int fsg_ctx_create(const fsg_ctx_config *cfg, fsg_ctx **out);
/* Waits for the context's stream and frees it.  Destroy the context's
 * serving engines and worlds first: they keep a pointer to it, so no call
 * on them (fsg_world_destroy included) is valid after this one. */
int fsg_ctx_destroy(fsg_ctx *ctx);
/* Bytes of world state currently held by the context. */
int fsg_ctx_usage(fsg_ctx *ctx, uint64_t *bytes_used);
/* Test hook: make the next n device allocations of this ctx fail. */
int fsg_ctx_inject_alloc_failures(fsg_ctx *ctx, int n);
/* Plain device-to-device copy of `bytes` bytes by a streaming kernel, timed
 * with HIP events (ms = mean over `reps`).  It has no counterpart in the
 * reference: it calibrates HBM counters and the achievable copy bandwidth. */
int fsg_ctx_copy_timed(fsg_ctx *ctx, size_t bytes, int reps, double *ms);
/* HIP events recorded on the ctx stream around whatever the caller enqueues
 * in between (device time of any call sequence on this context; no
 * reference counterpart).  stop waits for the second event. */
int fsg_ctx_timer_start(fsg_ctx *ctx);
int fsg_ctx_timer_stop(fsg_ctx *ctx, double *elapsed_ms);

/* ------------------------------------------------------------------------ */
/* World: replaces ME::world's grid and tick (world.hpp:63-193).             */
/* ------------------------------------------------------------------------ */

/* world::init (world.cpp:43-172): allocates the double-buffered device grid.
 * cfg may be NULL (fsg_default_config).  seed keys the counter RNG. */
int fsg_world_create(fsg_ctx *ctx, int width, int height, uint64_t seed,
                     const fsg_config *cfg, fsg_world **out);
int fsg_world_destroy(fsg_world *w);
int fsg_world_info(fsg_world *w, int *width, int *height, uint64_t *tick);
int fsg_world_set_tick(fsg_world *w, uint64_t tick); /* resume at tick */

/* Whole-grid host->device copy (the chunk load path writes real_tiles,
 * world.cpp:2653-2665).  n must equal width*height. */
int fsg_world_upload(fsg_world *w, const fsg_cell *cells, size_t n);
/* The benchmark world generator (fsg_generate) run on the device straight
 * into the world's grid: the same cells as fsg_generate + fsg_world_upload
 * without the host work (no reference counterpart: the reference fills its
 * grid from chunk files or its own generator, world.cpp:43-172, 2653-2665). */
int fsg_world_generate(fsg_world *w, int kind, uint64_t seed);
/* Rectangle device->host copy (getTile over a region, world.cpp:999-1002). */
int fsg_world_download(fsg_world *w, fsg_cell *out, int x0, int y0, int rw, int rh);

/* ------------------------------------------------------------------------ */
/* Chunk .pack images (Chunk::ChunkWrite chunk.cpp:219-330, Chunk::ChunkRead */
/* chunk.cpp:74-217), made and read on the device for a region of           */
/* ncx x ncy chunks of FSG_CHUNK^2 cells whose top-left cell is (x0, y0).    */
/* An image is the reference's file: i8 generation phase, i32 393216,        */
/* i32 compressed size, i32 65536, i32 compressed size, the LZ4 block of the */
/* 12-byte {u32 material index, u32 colour, i16 temperature, pad} records of */
/* tiles[] then layer2[], the LZ4 block of the u32 background.  Images are   */
/* ordered chunk row by chunk row.  The blocks decode with lz4's             */
/* LZ4_decompress_safe; their bytes are this library's own parse, not those  */
/* of LZ4_compress_fast.                                                     */
/* ------------------------------------------------------------------------ */
#define FSG_CHUNK 128 /* CHUNK_W = CHUNK_H (const.h:19-20) */

/* ChunkWrite for every chunk of the region: encodes the world's cells as
 * tiles[], layer2 (region-shaped host array, ncx*128 wide; NULL = empty
 * records) and background (likewise; NULL = 0), with generation phases
 * phase[ncx*ncy] (NULL = 0).  image_bytes[ncx*ncy] receives each image's
 * size; the images stay on the device until fsg_world_pack_download. */
int fsg_world_pack(fsg_world *w, int32_t x0, int32_t y0, int32_t ncx, int32_t ncy, const int8_t *phase,
                   const fsg_cell *layer2, const uint32_t *background, size_t *image_bytes);
/* Copies the images of the last fsg_world_pack, back to back, to out
 * (cap >= the sum of their sizes). */
int fsg_world_pack_download(fsg_world *w, void *out, size_t cap);
/* ChunkRead for every chunk of the region from images laid back to back in
 * host memory (image_bytes[ncx*ncy] their sizes), or, with images = NULL,
 * from the device-resident images of the last fsg_world_pack (same chunk
 * count; image_bytes ignored).  status[k] = FSG_OK, or FSG_ERR_INVALID for a
 * chunk whose header sizes, LZ4 blocks (a block that LZ4_decompress_safe
 * rejects or that decodes short) or material indices are bad; such a chunk
 * leaves its cells unchanged (ChunkRead logs the error and keeps what it
 * decoded, chunk.cpp:155-160).  phase, layer2 and background (region-shaped,
 * may be NULL) receive those parts of the images (zero layer2 and background
 * for a failed chunk; its phase is the header's when the header was read). */
int fsg_world_unpack(fsg_world *w, int32_t x0, int32_t y0, int32_t ncx, int32_t ncy, const void *images,
                     const size_t *image_bytes, int8_t *phase, fsg_cell *layer2, uint32_t *background,
                     int32_t *status);

/* Material index of the records.  ChunkWrite stores (u16)mat->id in the
 * record's index (chunk.cpp:241-242) and ChunkRead indexes materials_array
 * with it (chunk.cpp:167-177), so a file's indices are the reference's own
 * material ids (GENERIC_AIR 0, STONE 7, WATER 15, LAVA 16, OBSIDIAN 22,
 * STEAM 23: game_datastruct.cpp:72-110), not a world's.  to_file[m] is the
 * index written for material id m (< 256); from_file[i] is the material id
 * a record with index i reads as, or -1 for none (the chunk then fails, as
 * does any index >= 256).  MOVED (MaterialInstance::moved) has no record
 * field: it travels in bit 0 of the record's pad byte 10, which ChunkRead
 * ignores; FSG_PACK_NO_MOVED writes the pad as zero and reads every cell
 * with MOVED clear, as ChunkRead does.  The default map (and map = NULL) is
 * index = material id for ids < n_materials, MOVED kept; layer2 records
 * follow the same map.  Replaces materials_array / Material::id lookups of
 * chunk.cpp:167-177 and 241-242. */
#define FSG_PACK_NO_MOVED 1u
typedef struct fsg_pack_map {
    uint32_t to_file[16];
    int16_t from_file[256];
    uint32_t flags;
} fsg_pack_map;
/* Sets the map of the world's later fsg_world_pack / fsg_world_unpack calls
 * (copied; NULL restores the default).  FSG_ERR_INVALID if an id in use
 * (< n_materials) maps to an index >= 256 or a from_file entry names an id
 * >= n_materials. */
int fsg_world_set_pack_map(fsg_world *w, const fsg_pack_map *map);
/* The default map of a config (identity on its n_materials ids). */
int fsg_pack_map_default(const fsg_config *cfg, fsg_pack_map *map);

/* Decodes n raw LZ4 blocks on the device with the decode kernel that
 * fsg_world_unpack runs: block k is src[src_off[k], src_off[k] +
 * src_bytes[k]) and decodes into out[out_off[k], out_off[k] + cap[k]);
 * result[k] = the decoded size, or -1 for a block that LZ4_decompress_safe
 * (lz4.c:1949-2360, as chunk.cpp:151 and 191 call it) rejects.  Bytes of out
 * that no block writes keep their values.  A block check for callers and
 * tests; ChunkRead's blocks go through fsg_world_unpack. */
int fsg_lz4_decode_blocks(fsg_ctx *ctx, const void *src, size_t src_len, const int64_t *src_off,
                          const int32_t *src_bytes, const int32_t *cap, int32_t n, void *out, size_t out_len,
                          const int64_t *out_off, int32_t *result);

/* setTile (world.cpp:1004-1008) batched: edits are applied in array order
 * before the next tick; out-of-bounds edits are ignored like setTile's. */
typedef struct fsg_cell_edit {
    int32_t x, y;
    fsg_cell cell;
} fsg_cell_edit;
int fsg_world_set_cells(fsg_world *w, const fsg_cell_edit *edits, size_t n);
/* fsg_world_set_cells for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference: it is the same setTile batch
 * (same edits, same last-edit-wins result, same OOB rule), staged in a fixed
 * device array that every context on a GPU shares, instead of device blocks
 * allocated per call; it calls no hipMalloc, hipFree or other allocating
 * routine.  At most FSG_SET_CELLS_NOALLOC_MAX edits per call
 * (FSG_ERR_INVALID beyond).  Synchronous.  Calls on one GPU use the one
 * array and are serialised by a per-device lock inside the library (calls
 * from other processes have their own copy). */
#define FSG_SET_CELLS_NOALLOC_MAX 65536
int fsg_world_set_cells_noalloc(fsg_world *w, const fsg_cell_edit *edits, size_t n);

/* world::tick (world.cpp:1036-1948) ticks times, each followed by the
 * temperature step when the cadence says so (game.cpp:2157-2159).  Enqueued
 * on the ctx stream; returns when the launches are queued. */
int fsg_world_step(fsg_world *w, int ticks);
/* Same, synchronous, and reports elapsed device time measured with HIP
 * events on the ctx stream around the launches. */
int fsg_world_step_timed(fsg_world *w, int ticks, double *elapsed_ms);
/* world::tickTemperature (world.cpp:1950-2004): one standalone temperature
 * step for callers that drive the cadence themselves (temp_every = 0). */
int fsg_world_tick_temperature(fsg_world *w);
/* The tick zone, World::tickZone (world.hpp:138), which game.cpp:1629-1636
 * sets to the world less a one-chunk frame: only the rectangle
 * [x, x + width) x [y, y + height) ticks.  A Margolus block acts iff one of
 * its four cells lies in the zone, so cells leave the zone across its edges
 * (world.cpp:1073-1074 visits zone chunks, whose cells move anywhere) and a
 * cell of the one-cell ring outside moves only together with a zone cell.
 * Only zone cells take the temperature step (world.cpp:1953-1954), in the
 * tick and in fsg_world_tick_temperature, and fsg_world_particles_tick moves
 * only particles whose cell lies in the zone (world.cpp:2061).  The zone
 * must lie inside the world, width, height >= 1; a new world's zone, and
 * (0, 0, W, H), is the whole world.  The next tick processes every chunk. */
int fsg_world_set_tick_zone(fsg_world *w, int32_t x, int32_t y, int32_t width, int32_t height);
int fsg_world_get_tick_zone(fsg_world *w, int32_t *x, int32_t *y, int32_t *width, int32_t *height);
int fsg_world_sync(fsg_world *w);

/* Per-material counts and activity (the movingTiles[] census,
 * game.cpp:1991-2000, and the active chunk fraction). */
typedef struct fsg_stats {
    uint64_t tick;
    uint64_t count[16];       /* cells per material id                   */
    uint64_t checksum;        /* order-dependent 64-bit hash of the grid */
    uint64_t tiles_total;     /* 64x64 chunks in the world               */
    uint64_t tiles_processed; /* chunk updates run since the last reset  */
    uint64_t tiles_offered;   /* chunk slots offered since the last reset */
} fsg_stats;
int fsg_world_stats(fsg_world *w, fsg_stats *out);
int fsg_world_reset_activity(fsg_world *w);

/* ------------------------------------------------------------------------ */
/* Rigid-body pixels (game.cpp:1760-1815 stamp, 1896-1983 unstamp;           */
/* RigidBody::tiles, player.hpp:15-65).  Poses are fixed point, converted    */
/* from the host physics engine's float pose once per body per tick.         */
/* ------------------------------------------------------------------------ */
typedef struct fsg_body {
    int32_t pos_x_q16, pos_y_q16; /* b2Body::GetPosition, Q16.16 cells        */
    int32_t cos_q16, sin_q16;     /* cos/sin of b2Body::GetAngle, Q16.16      */
    int32_t mat_w, mat_h;         /* RigidBody::matWidth/matHeight            */
    uint32_t pixel_offset;        /* first pixel in the pixel array           */
    uint32_t enabled;             /* b2Body::IsEnabled                        */
} fsg_body;

typedef struct fsg_stamp_result {
    uint32_t placed;          /* pixels written into the grid               */
    uint32_t unplaced;        /* pixels with no free target in 5 offsets    */
    uint32_t displaced;       /* SAND/SOUP cells pushed out (CellData analogue) */
    uint32_t displaced_cap;   /* in: capacity of displaced_out              */
    fsg_cell_edit *displaced_out; /* in: optional host buffer for displaced cells */
    uint32_t *hits_sand;      /* in: optional [nbodies] damping counters    */
    uint32_t *hits_soup;      /* in: optional [nbodies]                     */
} fsg_stamp_result;

/* Stamps every enabled body's non-air pixels into the grid.  The world keeps
 * the bodies and pixels until fsg_world_unstamp_bodies. */
int fsg_world_stamp_bodies(fsg_world *w, const fsg_body *bodies, int nbodies,
                           const fsg_cell *pixels, size_t npixels,
                           fsg_stamp_result *res);
/* Pulls the pixels back out (same poses) and writes the updated pixel array
 * (npixels cells, same layout as stamped) to pixels_out. */
int fsg_world_unstamp_bodies(fsg_world *w, fsg_cell *pixels_out, size_t npixels);

/* ------------------------------------------------------------------------ */
/* Island check: world::physicsCheck (world.cpp:3330-3429) for a list of     */
/* seed cells.  The 4-connected SOLID component of each seed is searched up  */
/* to FSG_ISLAND_MAX_CELLS cells; a smaller one is cleared to air (colour 0, */
/* temperature 0: Tiles_NOTHING) and, above FSG_ISLAND_MIN_BODY cells, its   */
/* cells come back as the pixel image of a new rigid body (the reference's   */
/* makeRigidBody texture).  Seeds act in array order, as consecutive         */
/* physicsCheck calls would: a seed whose island an earlier seed removed     */
/* finds air.  The reference calls it once per tick at a random cell of the  */
/* tick zone (world.cpp:1929-1934) and from tools (game.cpp:700, 902-906).   */
/* ------------------------------------------------------------------------ */
#define FSG_ISLAND_MAX_CELLS 1000 /* larger components stay part of the world */
#define FSG_ISLAND_MIN_BODY 11    /* smaller islands are only deleted         */

typedef enum fsg_island_status {
    FSG_ISLAND_NONE = 0,     /* seed out of the world or not SOLID             */
    FSG_ISLAND_ATTACHED = 1, /* component larger than FSG_ISLAND_MAX_CELLS     */
    FSG_ISLAND_REMOVED = 2,  /* 1..10 cells, cleared                           */
    FSG_ISLAND_BODY = 3      /* 11..1000 cells, cleared, pixels returned       */
} fsg_island_status;

typedef struct fsg_island {
    int32_t x, y;          /* in: seed cell                                       */
    int32_t status;        /* out: fsg_island_status                              */
    uint32_t count;        /* out: cells of the island (REMOVED, BODY)            */
    int32_t min_x, min_y;  /* out: bounding box origin (REMOVED, BODY)            */
    int32_t w, h;          /* out: bounding box size                              */
    uint64_t pixel_offset; /* out: BODY: first cell of its w*h row-major image in
                            * pixels_out (air where the box is not the island)    */
} fsg_island;

/* Runs the check for n seeds (at most 2^20) and clears the islands it
 * removes.  The BODY images stay in device memory owned by the world until
 * the next check; *pixels_needed (may be NULL) gets their total cell count. */
int fsg_world_physics_check(fsg_world *w, fsg_island *seeds, int n, size_t *pixels_needed);
/* Copies the body images of the last check (pixel_offset indexes them);
 * FSG_ERR_INVALID if cap is smaller than their cell count. */
int fsg_world_island_pixels(fsg_world *w, fsg_cell *out, size_t cap);

/* ------------------------------------------------------------------------ */
/* Outlines of solid cells for the host physics engine: the marching-squares */
/* step of world::updateChunkMesh (world.cpp:722-895, which the mesh zone     */
/* runs for every chunk, world.cpp:961-998) and of updateRigidBodyHitbox      */
/* (world.cpp:288-510).  For each region the reference scans the cells in    */
/* row order for the next solid cell whose right, lower and lower-right     */
/* neighbours are not all solid and that no earlier loop marked (edgeSeen),  */
/* traces its perimeter with MarchingSquares::FindPerimeter                 */
/* (physics_math.cpp:1872-1965), marks the cells along it, simplifies the    */
/* loop's run-end vertices with Ramer-Douglas-Peucker at tolerance 1         */
/* (simplify, physics_math.cpp:1766-1812) and drops loops left with fewer    */
/* than 3 vertices; a loop that TPPLPoly::GetOrientation calls clockwise is  */
/* a hole.  The device does all of that; the host keeps TPPLPartition's      */
/* RemoveHoles / Triangulate_EC and the Box2D shapes.                        */
/* ------------------------------------------------------------------------ */
#define FSG_MESH_MAX_CELLS 16384 /* cells of one region (a 128x128 chunk) */
#define FSG_MESH_EDGE 1          /* region status: not meshed (see below) */

typedef struct fsg_mesh_region {
    int32_t x0, y0, w, h;        /* in: grid cells (world) or image size (islands) */
    int32_t status;              /* out: FSG_OK, FSG_MESH_EDGE, or FSG_ERR_INVALID */
    uint32_t first_loop, nloops; /* out: its loops in the loop array */
} fsg_mesh_region;

typedef struct fsg_mesh_loop {
    uint32_t first;           /* first vertex: (x, y) float pairs, region coordinates */
    uint32_t count;           /* vertices (>= 3), in the order of the reference's worldMesh;
                               * the reference fills its TPPLPoly in reverse */
    int32_t hole;             /* 1: clockwise by GetOrientation (SetHole(true)) */
    int32_t start_x, start_y; /* the cell the perimeter was traced from */
    uint32_t region;          /* index of its region */
} fsg_mesh_loop;

/* Outlines the SOLID cells of n regions of the grid (a region is meshed as
 * updateChunkMesh meshes a chunk: chunks are 128x128, at most
 * FSG_MESH_MAX_CELLS cells).  A region that reaches the world's last row or
 * column or lies outside it gets FSG_MESH_EDGE and no loops (the reference
 * skips chunks with chTx + CHUNK_W >= width, world.cpp:734-736).  The loops
 * stay in device memory owned by the world until the next mesh call;
 * *nloops / *nverts (may be NULL) get their totals. */
int fsg_world_mesh(fsg_world *w, fsg_mesh_region *regions, int n, uint32_t *nloops, uint32_t *nverts);
/* The same for the body images of the world's last island check
 * (fsg_world_physics_check): islands[i] with status FSG_ISLAND_BODY are
 * outlined over their w x h image (non-air pixels; the images are already
 * cropped to the island, as updateRigidBodyHitbox crops its surface);
 * regions[i] (n entries) receives each island's result, FSG_MESH_EDGE for
 * other statuses and for images 1 pixel wide or high (world.cpp:331, 365). */
int fsg_world_mesh_islands(fsg_world *w, const fsg_island *islands, int n, fsg_mesh_region *regions,
                           uint32_t *nloops, uint32_t *nverts);
/* Copies the loops and vertices (2 floats each) of the last mesh call. */
int fsg_world_mesh_download(fsg_world *w, fsg_mesh_loop *loops, uint32_t loop_cap, float *verts,
                            uint32_t vert_cap);

/* ------------------------------------------------------------------------ */
/* Free-fall particles: CellData (game_utils/cells.h:15-36) and              */
/* world::tickCells (world.cpp:2030-2189).  A particle carries one cell and  */
/* moves in f32 steps; when it enters a cell that is not AIR it lands: in    */
/* the cell it started the tick in if that is AIR, else in the first AIR     */
/* cell of a 33x33 square spiral around it; with none it is thrown up 16    */
/* cells.  Particles move in list order, and each sees the cells earlier     */
/* particles landed in during the same tick, as the reference's loop does.   */
/* The reference spawns them for cells a rigid body displaces (game.cpp:     */
/* 1788-1812) and for falling sand and liquid (world.cpp:1218-1225,          */
/* 1283-1305); this build's tick keeps falling cells on the grid, so the     */
/* body stamp's displaced cells are the source (fsg_stamp_result).           */
/* ------------------------------------------------------------------------ */
#define FSG_PARTICLE_PHASE 1u     /* CellData::phase: passes through cells     */
#define FSG_PARTICLE_TEMPORARY 2u /* CellData::temporary: dies at lifetime 0   */

typedef struct fsg_particle {
    float x, y, vx, vy, ax, ay;     /* CellData fields                          */
    float target_x, target_y, target_force;
    int32_t lifetime;
    uint32_t flags;                 /* FSG_PARTICLE_*                           */
    uint32_t in_object_state;       /* CellData::inObjectState (0..2)           */
    fsg_cell tile;                  /* the cell it carries                      */
} fsg_particle;

/* Appends n particles to the world's list (the reference's cells vector). */
int fsg_world_particles_add(fsg_world *w, const fsg_particle *p, size_t n);
/* One world::tickCells pass over the list, landing into the grid; the list
 * keeps its order, without the particles that landed, left the world or
 * ran out of lifetime.  *alive (may be NULL) gets the new count, *landed
 * (may be NULL) the cells written. */
int fsg_world_particles_tick(fsg_world *w, size_t *alive, size_t *landed);
/* Copies the list (at most cap particles) and its length. */
int fsg_world_particles_get(fsg_world *w, fsg_particle *out, size_t cap, size_t *n);

/* ------------------------------------------------------------------------ */
/* Texture pass (game.cpp:1994-2067): the grid as RGBA8 layers in device     */
/* memory, byte order R,G,B,A per texel, row-major with pitch width texels.  */
/* The reference rewrites the texels of dirty cells after every tick; every  */
/* texel it writes is a function of the cell alone (fire layer: of the cell  */
/* and the old texel), so one pass over a rectangle gives the same texture   */
/* without per-cell dirty flags.  Layers are allocated (zeroed: transparent) */
/* on first use and counted against the context's HBM budget.                */
/* ------------------------------------------------------------------------ */
#define FSG_LAYER_CELLS 1u    /* pixelsCells_ar: cell colour; AIR cells transparent 0 */
#define FSG_LAYER_EMISSION 2u /* pixelsEmission_ar: Material::emitColor; AIR 0        */
#define FSG_LAYER_FIRE 4u     /* pixelsFire_ar: fire cells' colour; AIR 0; others kept */

typedef struct fsg_render_params {
    uint32_t layers;         /* FSG_LAYER_* mask                                      */
    int32_t x0, y0, w, h;    /* rectangle, clipped to the world; w or h <= 0: whole world */
    uint32_t fire_mat;       /* material id drawn on the fire layer (FIRE); 255: none  */
    uint32_t count_moved;    /* 1: fill fsg_render_stats (reads the state before the last tick) */
    uint32_t emit[16];       /* Material::emitColor per material id, 0xAARRGGBB        */
} fsg_render_params;

typedef struct fsg_render_stats {
    uint64_t moved[16];      /* movingTiles (game.cpp:1999): cells in the rectangle whose
                              * material or colour changed in the last tick, by new material */
    uint64_t fire_cells;     /* fire cells drawn (hadFire)                             */
    uint32_t had_dirty;      /* any moved cell (hadDirty)                              */
    uint32_t reserved;
} fsg_render_stats;

/* Updates the requested layers over the rectangle.  Enqueued on the ctx
 * stream; with stats != NULL it waits and fills *stats (all zero if the
 * world has not ticked since its last upload). */
int fsg_world_render(fsg_world *w, const fsg_render_params *p, fsg_render_stats *stats);
/* Device address of a layer (NULL until first rendered); pitch = width texels. */
int fsg_world_layer(fsg_world *w, uint32_t layer, uint32_t **device_ptr);
/* Rectangle device->host copy of a layer (R_UpdateImageBytes' source). */
int fsg_world_layer_download(fsg_world *w, uint32_t layer, int x0, int y0, int rw, int rh,
                             uint32_t *out);

/* ------------------------------------------------------------------------ */
/* Batching: many worlds, one launch per tick over a multi-world chunk list, */
/* ticks captured as HIP graphs.  Results are bit-identical to stepping each */
/* world alone.                                                              */
/* ------------------------------------------------------------------------ */
int fsg_batch_step(fsg_ctx *ctx, fsg_world *const *worlds, int n, int ticks);
int fsg_batch_step_timed(fsg_ctx *ctx, fsg_world *const *worlds, int n, int ticks,
                         double *elapsed_ms);
/* Per-world launches of the same work, for comparison. */
int fsg_perworld_step_timed(fsg_ctx *ctx, fsg_world *const *worlds, int n, int ticks,
                            double *elapsed_ms);

/* ------------------------------------------------------------------------ */
/* Autotuner for batched launches (no reference counterpart: the reference  */
/* ticks one world on a CPU thread pool).  Every choice gives bit-identical  */
/* results; only the speed differs.                                          */
/* ------------------------------------------------------------------------ */
typedef struct fsg_tuning {
    int32_t graph_ticks; /* ticks per captured HIP graph block: 1 (no graph), 4, 8, 16 or 32 */
    int32_t chunk_skip;  /* dirty-chunk skipping in batches: -1 each world's own setting, 0 off, 1 on */
    int32_t occupancy;   /* tick-kernel register budget: 4 or 5 workgroups per CU */
    int32_t reserved;    /* 0 */
} fsg_tuning;

#define FSG_TUNE_MAX_CANDIDATES 32
typedef struct fsg_tune_report {
    fsg_tuning best;
    int32_t n_candidates;  /* candidates timed (0 when the choice came from the cache) */
    int32_t from_cache;    /* 1: the cache held a choice for this GPU and world-size class */
    uint64_t size_class;   /* the cache key's world-size signature */
    fsg_tuning candidate[FSG_TUNE_MAX_CANDIDATES];
    double ms[FSG_TUNE_MAX_CANDIDATES]; /* device time of sample_ticks ticks per candidate */
} fsg_tune_report;

/* Current batch tuning of the context (default {16, -1, 5, 0}). */
int fsg_ctx_get_tuning(fsg_ctx *ctx, fsg_tuning *out);
int fsg_ctx_set_tuning(fsg_ctx *ctx, const fsg_tuning *t);
/* Times every candidate tuning on copies of the given worlds (the worlds
 * themselves do not advance), sample_ticks ticks each, and makes the fastest
 * the context's tuning.  Choices are cached per GPU architecture and
 * world-size class (a histogram of log2 world areas), so a later call on a
 * similar batch returns at once. */
int fsg_autotune_batch(fsg_ctx *ctx, fsg_world *const *worlds, int n, int sample_ticks,
                       fsg_tune_report *report);

/* ------------------------------------------------------------------------ */
/* Serving: a request loop that keeps worlds resident in HBM and steps them  */
/* in continuous batches.  The reference runs one world at a fixed 30 TPS    */
/* from game::run (game.cpp:1009-1018); there is no request API to mirror,   */
/* so these entry points have no reference counterpart.                      */
/*                                                                           */
/* A request names a world by a caller-chosen id.  An unknown id creates the */
/* world (fsg_generate kind/size/seed, plus n_bodies rigid bodies whose      */
/* poses the server steps on the host and stamps/unstamps every tick); a     */
/* known id steps the resident world.  Requests are admitted while the HBM   */
/* state budget of the context allows (idle resident worlds are evicted,     */
/* least recently used first); all admitted requests advance together in one */
/* batched launch per step, and new requests join between steps.  A world's */
/* result does not depend on which other requests share its batch.          */
/* ------------------------------------------------------------------------ */
typedef struct fsg_server fsg_server;

typedef struct fsg_serve_config {
    int32_t max_active;      /* requests stepped together (0 = 1024)            */
    int32_t block_ticks;     /* ticks per step (at most; 0 = 16)                */
    int32_t verify_resident; /* 1: check a resident world's checksum on reuse   */
    int32_t gen_threads;     /* host threads generating new worlds (0 = 4)      */
} fsg_serve_config;

typedef struct fsg_serve_request {
    uint64_t world_id;
    int32_t kind;            /* fsg_generate kind for a new world               */
    int32_t width, height;   /* new world size                                  */
    int32_t ticks;           /* ticks to advance (>= 1)                         */
    uint64_t seed;           /* new world seed                                  */
    int32_t n_bodies;        /* new world: rigid bodies, 0..4096                */
    int32_t reserved;
    uint64_t deadline_ns;    /* fsg_now_ns() deadline; 0 = none                 */
} fsg_serve_request;

typedef struct fsg_serve_result {
    uint64_t request_id, world_id;
    int32_t status;          /* FSG_OK, FSG_ERR_BUDGET/NOMEM/TIMEOUT/CANCELLED/CHECKSUM */
    int32_t ticks_done;
    uint64_t tick;           /* world tick after the request                    */
    uint64_t checksum;       /* grid checksum after the request (fsg_stats)     */
    uint64_t cells;          /* cells of the world                              */
    uint64_t t_submit_ns, t_admit_ns, t_done_ns;
} fsg_serve_result;

typedef struct fsg_serve_stats {
    uint64_t submitted, completed, failed;
    uint64_t queued, active, resident_worlds, resident_bytes;
    uint64_t steps, cell_updates, evictions;
} fsg_serve_stats;

uint64_t fsg_now_ns(void);
int fsg_serve_create(fsg_ctx *ctx, const fsg_serve_config *cfg, fsg_server **out);
int fsg_serve_destroy(fsg_server *s);
int fsg_serve_submit(fsg_server *s, const fsg_serve_request *req, uint64_t *request_id);
int fsg_serve_cancel(fsg_server *s, uint64_t request_id);
/* Completed results in completion order; *n = number written (<= cap). */
int fsg_serve_poll(fsg_server *s, fsg_serve_result *out, int cap, int *n);
/* Blocks until every submitted request has completed or timeout_ns passed. */
int fsg_serve_drain(fsg_server *s, uint64_t timeout_ns);
int fsg_serve_release(fsg_server *s, uint64_t world_id);
int fsg_serve_stats_get(fsg_server *s, fsg_serve_stats *out);
/* Test hook: overwrite one cell of an idle resident world, so that the next
 * request on it fails the resident checksum check (verify_resident = 1). */
int fsg_serve_corrupt(fsg_server *s, uint64_t world_id);

/* ------------------------------------------------------------------------ */
/* Host-side generator of the benchmark worlds (not on the device path).     */
/* kind: 1 = basin terrain + sand discs (512^2 world), 2 = kind 1 + lava and  */
/* oil discs, 3 = all-active striped world, 4 = kind 2 with disc counts       */
/* scaled to the world's area.                                               */
/* ------------------------------------------------------------------------ */
int fsg_generate(int kind, int width, int height, uint64_t seed, fsg_cell *out);

/* Test hook with no counterpart in the reference: evaluates the device
 * block rule compiled for the host on one block q[4] = TL, TR, BL, BR at
 * iteration it with random words r, s.  form 0 is the staged rule, form 1
 * the table-driven form the tick kernel runs (reacting blocks routed to the
 * staged rule, as in the kernel).  Returns 1 if the block changed, 0 if not,
 * or a negative status. */
int fsg_rule_eval(const fsg_config *cfg, int it, uint32_t r, uint32_t s, uint64_t q[4], int form);

/* splitmix64 output for state x (Vigna), and the per-phase RNG key. */
uint64_t fsg_splitmix64(uint64_t x);
uint64_t fsg_phase_key(uint64_t seed, uint64_t phase_index);

#ifdef __cplusplus
}
#endif
#endif /* FSG_H */
