/*
 * powerfoam_amd.h -- C-ABI of the MI355X-native Power Foam renderer hot path.
 *
 * Plain C: raw pointers, sizes and an opaque stream handle. No HIP, torch or
 * C++ types appear here, so a gcc-compiled client (ctypes, cgo, JNI) can bind
 * it. Every device pointer is a HIP device address; `stream` is a hipStream_t
 * passed as void* (NULL = the default stream). No entry point allocates or
 * frees device memory: callers own every buffer, and scratch space is passed
 * in explicitly with its size given by the matching *_workspace_size query.
 *
 * Every entry point returns PF_OK (0) or a negative pf_status; a readable
 * message for the last failure on the calling thread is in pf_last_error().
 *
 * Conventions shared with the reference (theialab/powerfoam):
 *   - screen tiles are 8x8 pixels, one tile = 64 rays (rasterize.py:10-12);
 *   - tile index is column-major: tile = tile_i + tile_j * tiles_h, and the
 *     pixel inside a tile is column-major too: i = lane % 8, j = lane / 8
 *     (rasterize.py:919-938);
 *   - tiles_h = 1 + (H-1)/8, tiles_w = 1 + (W-1)/8.
 */
#ifndef POWERFOAM_AMD_H
#define POWERFOAM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_TILE_WIDTH 8
#define PF_TILE_SIZE 64
#define PF_MAX_TEXELS 8 /* texture.py:7 max_sites */

typedef enum pf_status {
  PF_OK = 0,
  PF_ERR_ARG = -1,         /* a null/inconsistent argument */
  PF_ERR_OVERFLOW = -2,    /* an output or workspace is too small */
  PF_ERR_HIP = -3,         /* a HIP runtime call failed */
  PF_ERR_UNSUPPORTED = -4  /* a mode the build does not provide */
} pf_status;

/* Camera: the reference's WarpCamera (camera.py:8-16). `ray_maps` is an
 * [H,W,6] float array (origin xyz, direction xyz) for non-pinhole cameras, or
 * NULL for a pinhole camera whose rays come from eye/right/up
 * (camera.py:19-24). */
typedef struct pf_camera {
  float eye[3];
  float right[3];
  float up[3];
  int32_t width;
  int32_t height;
  const float* ray_maps;
  float fov_cos_cutoff;
} pf_camera;

/* Per-site scene records (rasterize.py:1619-1624 packing). */
typedef struct pf_scene {
  int32_t n_sites;                  /* N */
  int32_t n_texels;                 /* T, 1..8 */
  const float* spheres;             /* [N,4]  centre xyz, radius */
  const float* nsigmas;             /* [N,4]  normal xyz, density */
  const float* texel_sites;         /* [N,T,3] world-space texel sites */
  const float* texel_rgbh;          /* [N,T,4] texel rgb, height */
  const int32_t* adjacency;         /* [E]   neighbour site ids */
  /* [N+1] CSR offsets into adjacency; the kernels trust them:
   * adjacency_offsets[N] must equal E, the length of adjacency and
   * adjacency_diff (the Python mirror checks it at the frame's host sync) */
  const int32_t* adjacency_offsets;
  const uint16_t* adjacency_diff;   /* [E,4] IEEE binary16 (c_adj - c, r_adj) */
  /* [N,T,4] bfloat16 (rgb, height), 8-byte aligned, or NULL. When set, the
   * rasterizer kernels (pf_raster_fwd/bwd/rgb/vis) read the texel features
   * from it instead of texel_rgbh (which may then be NULL): the features
   * stored in BF16, geometry and transmittance still FP32. Not in the
   * reference (its fp16 attributes, rasterize.py:2008-2014, are the
   * benchmark pass's); the ray tracer reads texel_rgbh. */
  const uint16_t* texel_rgbh_bf16;
  /* [N] or NULL: the compositing (pf_raster_fwd / pf_raster_bwd) reads only
   * the first adjacency_count[s] faces of each site's range (the kept faces
   * pf_prune_faces puts first); NULL: every face. Other entry points read the
   * whole range. Not in the reference. */
  const int32_t* adjacency_count;
} pf_scene;

/* Depth-sorted per-tile site lists produced by binning. */
typedef struct pf_tile_lists {
  int32_t n_tiles;
  const int64_t* offsets; /* [n_tiles+1] */
  const int32_t* sites;   /* [M] */
  /* [n_tiles] or NULL: the order in which tiles are started (a permutation).
   * Results do not depend on it; starting the longest lists first shortens
   * the kernels' tail. NULL: natural order, dealt to the XCDs in runs. */
  const int32_t* order;
} pf_tile_lists;

/* Optional splitting of long tile lists (no reference counterpart; the
 * results are those of the unsplit walk up to float summation order). A list
 * of `count` sites becomes max(1, ceil(count / seg_len)) segments that run in
 * parallel: the forward first sums -sigma dt over every segment but a tile's
 * last (log_t_sums), each segment then starts from the exact transmittance
 * the earlier ones leave; the backward starts each segment from the adjoint
 * state the later ones leave. n_segs = 0 keeps one wave per whole list.
 * The forward and the backward of a frame must use the same table. */
typedef struct pf_segments {
  int32_t n_segs;             /* table capacity (>= n_tiles) */
  const int32_t* count;       /* device int: segments in use (<= n_segs), or NULL: all n_segs */
  int32_t seg_len;            /* sites per segment, a positive multiple of 64 */
  const int32_t* tile_first;  /* [n_tiles+1] first segment id of each tile */
  const int32_t* seg_tile;    /* [n_segs] tile of each segment id */
  const int32_t* order;       /* [n_segs] segment ids in start order (longest first) */
  float* log_t_sums;          /* [n_segs,64] written by the forward, read by the backward */
  uint8_t* lane_active;       /* [n_segs,64] written by the forward, read by the backward */
  float* scratch;             /* [n_segs,64] per-call scratch */
} pf_segments;

typedef struct pf_raster_fwd_args {
  pf_camera cam;
  pf_scene scene;
  pf_tile_lists tiles;
  pf_segments segs;
  const float* depth_quantiles; /* [H,W,Q] transmittance levels, descending, or NULL */
  int32_t n_quantiles;          /* Q (0 when depth_quantiles is NULL) */
  const float* ray_gt;          /* [H,W,3] or NULL */
  float transmittance_threshold;
  int32_t return_point_err;
  int32_t loss_accumulation;    /* render_objective == "surface" */
  /* outputs (color, normal_distance, normal, err, contrib and point_err
   * must be zeroed and quantile_depths set to -1 by the caller, as the
   * reference allocates them) */
  int32_t* tile_early_stop; /* [n_tiles] */
  float* color;             /* [H,W,3] */
  float* log_t;             /* [H,W]   */
  float* normal_distance;   /* [H,W]   */
  float* normal;            /* [H,W,3] */
  float* quantile_depths;   /* [H,W,Q] or NULL */
  float* err;               /* [H,W] or NULL (loss_accumulation only) */
  float* contrib;           /* [N] accumulated */
  float* point_err;         /* [N] accumulated, or NULL */
  /* [M] or NULL, zeroed by the caller: set to 1 for each list entry whose
   * site shades some ray of its tile (an interval of positive length).
   * Not in the reference: the backward skips the other entries. */
  uint8_t* entry_shades;
  /* The learned colour head evaluated inside the compositing (pf_mlp_head,
   * below), or NULL: texel colours come from scene.texel_rgbh. Not in the
   * reference. */
  const struct pf_mlp_head* mlp_head;
} pf_raster_fwd_args;

typedef struct pf_raster_bwd_args {
  pf_camera cam;
  pf_scene scene;
  pf_tile_lists tiles;
  pf_segments segs;
  const float* depth_quantiles; /* [H,W,Q] or NULL */
  int32_t n_quantiles;
  const int32_t* tile_early_stop;      /* [n_tiles] from the forward */
  const float* log_t;                  /* [H,W] final log-transmittance */
  const float* grad_color;             /* [H,W,3] */
  const float* grad_log_t;             /* [H,W] = -grad_opacity * exp(log_t) */
  const float* grad_normal_distance;   /* [H,W] */
  const float* grad_normal;            /* [H,W,3] */
  const float* grad_quantile_depths;   /* [H,W,Q] or NULL */
  const float* grad_contrib;           /* [N] */
  /* 0: the reference's backward, term for term (rasterize.py:1150-1514).
   * 1: the exact derivative of the forward: no 1e-6 regularisers, no texel
   *    weight cut-off, the far-end surface clip routed to the plane only, and
   *    the normal_distance term differentiated as (n.d)^2. Used by the
   *    finite-difference checks. */
  int32_t exact_gradients;
  /* outputs, accumulated (zero them first) */
  float* grad_spheres;      /* [N,4] */
  float* grad_nsigmas;      /* [N,4] */
  float* grad_texel_sites;  /* [N,T,3] */
  float* grad_texel_rgb;    /* [N,T,3] */
  float* grad_texel_height; /* [N,T] */
  /* [M] from the forward (pf_raster_fwd_args.entry_shades), or NULL: the
   * walk skips entries marked 0, which add nothing to any gradient */
  const uint8_t* entry_shades;
} pf_raster_bwd_args;

/* Benchmark forward: colour only. scene.adjacency_diff holds the power-face
 * encoding of benchmark.py:297-302, binary16 (c_adj - c, pm_adj - pm) with
 * pm = (|c|^2 - r^2)/2. */
typedef struct pf_raster_rgb_args {
  pf_camera cam;
  pf_scene scene;
  pf_tile_lists tiles;
  float transmittance_threshold;
  float* color; /* [H,W,3], every in-image pixel written */
} pf_raster_rgb_args;

/* Visualisation pass (visualization_kernel, rasterize.py:739-890; VisOptions
 * :15-20): like pf_raster_rgb, but faces from the training encoding
 * (c_adj - c, r_adj) of pf_adj_encode, a cap on the list entries each ray
 * looks at, and extra outputs. */
typedef struct pf_raster_vis_args {
  pf_camera cam;
  pf_scene scene;
  pf_tile_lists tiles;
  float transmittance_threshold;
  int32_t max_intersections;
  float depth_quantile;
  float bkgd_color[3];
  float* color;           /* [H,W,3]: composited colour + bkgd * T */
  float* depth;           /* [H,W]: where T first drops below depth_quantile, 0 if never */
  float* normal;          /* [H,W,3]: sum of alpha T n */
  float* alpha;           /* [H,W]: 1 - T */
  int32_t* intersections; /* [H,W]: list entries looked at (the stopping one included) */
} pf_raster_vis_args;

/* Ray tracer (raytrace.py:115-241). scene.adjacency must be the full power
 * (weighted Delaunay) adjacency; scene.adjacency_diff holds the power-face
 * encoding (as pf_raster_rgb_args) or is NULL for exact geometry computed from
 * the spheres. Rays start in start_sites[pixel] when given, else start_site. */
typedef struct pf_trace_args {
  pf_camera cam;
  pf_scene scene;
  int32_t start_site;
  const int32_t* start_sites;   /* [H*W] or NULL */
  float transmittance_threshold;
  float depth_quantile;         /* depth output: where T first drops below this */
  int32_t max_steps;            /* walk length cap per ray (<= 0: 65536) */
  /* outputs */
  float* color;    /* [H,W,3] */
  float* log_t;    /* [H,W]   or NULL */
  float* normal;   /* [H,W,3] or NULL: sum of alpha T n */
  float* depth;    /* [H,W]   or NULL: -1 where T never drops below depth_quantile */
  int32_t* q_step; /* [H,W]   or NULL: walk step of the depth crossing (backward input) */
  float* q_sigma;  /* [H,W]   or NULL: density of that cell (backward input) */
  int32_t* cells;  /* [H,W,max_cells] or NULL: visited cells in order */
  int32_t max_cells;
  int32_t* n_cells;/* [H,W] or NULL: number of cells visited */
  /* > 0: ballot compaction over groups of `compact` neighbouring tiles: a
   * wave walks the group's rays and hands a lane whose ray has ended the
   * group's next unstarted ray (its slot from a ballot of the lanes that
   * ended in the same step); 0: one wave per tile, a lane idles once its ray
   * ends. Same results either way. */
  int32_t compact;
  /* [n_tiles] or NULL (compact = 0 only): per 8x8 tile, the longest walk
   * (cells visited) of its rays, for pf_trace_tile_order. n_tiles =
   * ceil(H/8) * ceil(W/8), tile t = column-major (t % ceil(H/8) the tile row). */
  int32_t* tile_steps;
  /* [E,4] or NULL, read when scene.adjacency_diff is NULL: the exact face
   * planes from pf_trace_planes for the same spheres and adjacency. The walk
   * then reads one contiguous record per face instead of the neighbour's id
   * and sphere; same values, same results bit for bit. */
  const float* face_planes;
} pf_trace_args;

/* Exact adjoint of pf_trace_fwd (no reference counterpart: the reference ray
 * tracer has no backward). Re-walks each ray front to back. */
typedef struct pf_trace_bwd_args {
  pf_trace_args fwd;           /* same inputs as the forward; its output pointers are ignored */
  const float* color;          /* [H,W,3] saved forward outputs */
  const float* normal;         /* [H,W,3] */
  const float* log_t;          /* [H,W] */
  const int32_t* q_step;       /* [H,W] */
  const float* q_sigma;        /* [H,W] */
  const float* grad_color;     /* [H,W,3] */
  const float* grad_normal;    /* [H,W,3] */
  const float* grad_depth;     /* [H,W] or NULL */
  const float* grad_opacity;   /* [H,W] */
  float* grad_spheres;         /* [N,4] accumulated */
  float* grad_nsigmas;         /* [N,4] */
  float* grad_texel_sites;     /* [N,T,3] */
  float* grad_texel_rgb;       /* [N,T,3] */
  float* grad_texel_height;    /* [N,T] */
  /* [n_tiles] or NULL: the order in which the backward's waves take the tiles
   * (a permutation of 0..n_tiles-1, e.g. from pf_trace_tile_order); NULL:
   * tile index order. The summation order of the gradients follows it. */
  const int32_t* tile_order;
} pf_trace_bwd_args;

/* ---- library ---------------------------------------------------------- */
int pf_version(void);
const char* pf_last_error(void);

/* ---- binning (replaces count_visible_pinhole_kernel rasterize.py:140-200,
 *      torch.cumsum + .item() rasterize.py:1713-1714,
 *      write_visible_pinhole_kernel rasterize.py:293-363,
 *      torch.argsort + gather rasterize.py:1761-1768) --------------------- */

/* Zeroes tile_counts[0..n_tiles] and counts, per tile, the sites whose
 * projected footprint passes the tile test; tile t's count lands in
 * tile_counts[t+1] so an inclusive scan gives offsets directly. `live`
 * (uint8 [n_sites], from pf_live_sites) leaves out the sites marked 0; NULL
 * counts every site, as the reference does. `bounds` ([n_sites,4] f32 balls
 * from pf_live_sites, 16-byte aligned, or NULL) keeps a (tile, site) pair
 * only when the tile also meets the site's ball footprint. */
int pf_bin_count(const pf_camera* cam, const float* spheres, const uint8_t* live,
                 const float* bounds, int32_t n_sites, int32_t* tile_counts, void* stream);

/* offsets[k] = sum(tile_counts[0..k]) as int64 (torch.cumsum of the counts). */
int pf_bin_scan(const int32_t* tile_counts, int32_t n_tiles, int64_t* offsets,
                void* stream);

/* Writes one 64-bit record per (tile, site) pair into the tile's segment:
 * (bits(|c-eye|^2 - r^2) << 32) | site. `cursor` is an int32 [n_tiles]
 * scratch array the call zeroes. Sets visible[site] = 1 for every site with
 * a tile (the reference's prim_visible_mask, rasterize.py:1761-1765), also
 * for sites that `live` leaves out of the lists; the caller zeroes `visible`
 * first (or passes NULL). `live` and `bounds` as for pf_bin_count, and the
 * same arrays. A record whose slot lies past its tile's segment (offsets
 * that do not come from the same count) is not written, as the reference's
 * guard `idx < offsets[tile+1]` (rasterize.py:358) drops it; the call adds
 * the number of such records to *dropped (a device int32 the caller zeroes,
 * or NULL), so a host can report the overflow at its next synchronisation. */
int pf_bin_emit(const pf_camera* cam, const float* spheres, const uint8_t* live,
                const float* bounds, int32_t n_sites, const int64_t* offsets, int32_t n_tiles,
                uint64_t* records, int32_t* cursor, uint8_t* visible, int32_t* dropped, void* stream);

/* Face encodings pf_live_sites reads. */
enum {
  PF_FACES_RADIUS = 0, /* (c_adj - c, r_adj) from pf_adj_encode */
  PF_FACES_POWER = 1   /* (c_adj - c, pm_adj - pm), pm = (|c|^2 - r^2)/2 (benchmark.py:297-302) */
};

/* live[s] = 0 for a site whose sphere lies wholly beyond the plane of one of
 * its power faces, by more than the rounding margin of the compositing
 * kernels for rays from cam->eye; 1 otherwise. Such a site has an empty
 * clipped interval on every ray, so the compositing results (colour,
 * transmittance, normals, depths, contributions, every gradient) are the
 * same whether or not it is binned. The faces are read as the compositing
 * kernels read them: `diff` in the form `diff_kind` names, or, with diff
 * NULL, the full-precision spheres of `adjacency`. Pinhole cameras only.
 * With `bounds` ([n_sites,4] f32, 16-byte aligned, or NULL) it also writes
 * per site a ball (centre, radius) that holds every point the site can
 * shade: the sphere (inflated for FP32 rounding), shrunk to the enclosing
 * ball of its cap on the cell's side of each face that leaves the ball's
 * centre outside, three sweeps over the faces; a site whose ball falls
 * wholly outside a face gets live 0 as well. n_edges is the length of the
 * adjacency (and diff); offsets beyond it end a site's faces there.
 * Not in the reference: the reference bins every site in the footprint
 * (count_visible_pinhole_kernel rasterize.py:140-200 has no such test). */
int pf_live_sites(const pf_camera* cam, const float* spheres, int32_t n_sites,
                  const int32_t* adjacency, const int32_t* adjacency_offsets,
                  int64_t n_edges, const uint16_t* diff, int32_t diff_kind, uint8_t* live,
                  float* bounds, void* stream);

/* Ray-map cameras (cam->ray_maps set; rasterize.py:1681-1760 when
 * args.is_pinhole is false) bin through a quadtree of ray cones instead of
 * the projected ellipse.
 *
 * Number of cone nodes (float4 each) for a width x height image, or -1 when
 * the size is outside 5..8192 pixels a side. Replaces the node count of
 * _precompute_cones (rasterize.py:1518-1550). Host-only, no GPU work. */
int32_t pf_cone_tree_nodes(int32_t width, int32_t height);

/* Builds the cone tree of cam->ray_maps into cones [nodes, 4] (axis, cos of
 * the half-angle; cos > 1 = no valid ray). Replaces
 * compute_leaf_cones_kernel + merge_cones_kernel launches
 * (rasterize.py:468-569, 1552-1583). */
int pf_bin_cones(const pf_camera* cam, float* cones, void* stream);

/* pf_bin_count for ray-map cameras: count_visible_generic_kernel
 * (rasterize.py:202-291). */
int pf_bin_count_cones(const pf_camera* cam, const float* cones, const float* spheres, int32_t n_sites,
                       int32_t* tile_counts, void* stream);

/* pf_bin_emit for ray-map cameras: write_visible_generic_kernel
 * (rasterize.py:365-466, guard :429). Records and their order after
 * pf_bin_sort, and *dropped, are as for pf_bin_emit. */
int pf_bin_emit_cones(const pf_camera* cam, const float* cones, const float* spheres, int32_t n_sites,
                      const int64_t* offsets, int32_t n_tiles, uint64_t* records, int32_t* cursor,
                      uint8_t* visible, int32_t* dropped, void* stream);

/* Bytes of scratch pf_bin_sort can use for these tile segments. */
size_t pf_bin_sort_workspace_size(int64_t n_records, int32_t n_tiles);

/* Sorts every tile segment ascending by record (depth key, then site id) and
 * writes the site ids to `sites` [M]. Deterministic. The workspace is
 * optional: with it, segments longer than 4,096 records are listed and sorted
 * by 1,024 threads each with 16,384 records in LDS; without it they merge
 * through global memory in the one-workgroup-per-tile kernel. */
int pf_bin_sort(const int64_t* offsets, int32_t n_tiles, uint64_t* records,
                int32_t* sites, void* workspace, size_t workspace_bytes,
                void* stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): pf_bin_sort with its workspace (the list of long tiles) in a
 * library-owned device array of PF_SORT_STATIC_MAX_TILES + 1 int32 per GPU
 * (1 MiB + 4 B, held until the process ends), so the caller passes inputs
 * and outputs only. Same results as pf_bin_sort with a workspace. The first
 * call on a device looks up that device's array address (no allocation).
 * Returns PF_ERR_OVERFLOW above PF_SORT_STATIC_MAX_TILES tiles (a 4096 x 4096
 * image). Every call on one GPU shares its array: calls that overlap on one
 * GPU (concurrent streams or threads) race; calls in one stream order are
 * safe, as nothing in the array must start at zero. Allocates nothing. */
#define PF_SORT_STATIC_MAX_TILES 262144
int pf_bin_sort_static(const int64_t* offsets, int32_t n_tiles, uint64_t* records,
                       int32_t* sites, void* stream);

/* Dispatch order of the compositing kernels, built on the device from the
 * tile offsets (no reference counterpart: the reference starts tiles in index
 * order; this replaces host-side argsorts and a host sync). Writes
 *   tile_order [n_tiles]: tiles by descending list length (pf_tile_lists.order);
 *   tile_first [n_tiles+1], seg_tile, seg_order [max_segs], *count: the
 *   segment table of pf_segments for segments of seg_len sites, segments by
 *   descending size.
 * Orders are by length bucket: exact below 512 sites, in steps of 32 sites
 * up to 16,864, one bucket above. n_records = offsets[n_tiles]; the table
 * needs max_segs >= n_tiles + n_records / seg_len; seg_len is a positive
 * multiple of 64 up to 1024. tile_order, or the four segment outputs, may be
 * NULL to skip that part. */
int pf_tile_schedule(const int64_t* offsets, int32_t n_tiles, int64_t n_records, int32_t seg_len,
                     int32_t max_segs, int32_t* tile_order, int32_t* tile_first, int32_t* seg_tile,
                     int32_t* seg_order, int32_t* count, void* stream);

/* ---- Cech adjacency (AABBTree.update / build_cech_complex bvh.py:76-123,
 *      count_adjacent / write_adjacent bvh.py:7-69) ----------------------
 * j != i is a neighbour of site i when the float32 boxes c -+ r overlap and
 * |c_i - c_j| < r_i + r_j, with c and r taken back from the box as the
 * reference does. Sequence (one stream, the workspace untouched between):
 *   pf_cech_count  -> counts [N+1] (counts[0] = 0, counts[i+1] = degree of i)
 *   host: offsets = inclusive cumsum of counts (int64 [N+1]), M = offsets[N]
 *   pf_cech_emit   -> neighbour ids as records [M] in each site's segment
 *   pf_bin_sort(offsets, N, records, adjacency, ...) -> adjacency [M] int32,
 *                     ids ascending within every site (the reference's BVH
 *                     order is unspecified).
 * spheres [N,4] = (x, y, z, r), 16-byte aligned, N <= PF_CECH_MAX_SITES; the
 * workspace is pf_cech_workspace_size(N) bytes (about 48 N + 12 MiB at most
 * for the hash table), 256-byte aligned. Entries beyond a segment are added
 * to *dropped (cleared by the caller), never written. Non-finite centres or
 * radii make no pairs of their own and never stop the build. */
#define PF_CECH_MAX_SITES (1 << 29)
size_t pf_cech_workspace_size(int32_t n_sites);
int pf_cech_count(const float* spheres, int32_t n_sites, void* workspace, size_t workspace_bytes,
                  int32_t* counts, void* stream);
int pf_cech_emit(const float* spheres, int32_t n_sites, void* workspace, size_t workspace_bytes,
                 const int64_t* offsets, uint64_t* records, int32_t* dropped, void* stream);

/* ---- restricted power adjacency ------------------------------------------
 * No reference kernel: replaces the lifted-hull power adjacency with the
 * overlap filter (build_power_adjacency, benchmark.py:25-83) for rendering.
 * From symmetric, ascending Cech lists (pf_cech_*): keeps the entries (a, b)
 * whose power face meets the two balls (FP64 disk-and-half-planes test on the
 * radical plane against a's other neighbours; both directed entries take the
 * answer of a < b), which is every face that bounds a site's cell inside its
 * ball. A site left with none whose cell misses its ball keeps, one way, the
 * neighbours whose radical planes cut into its ball, so it still renders
 * empty. keep [n_edges] (0/1) and counts [N+1] (counts[0] = 0, kept entries
 * per site) are overwritten; the lists are adjacency[keep] with offsets the
 * cumulative counts. Neighbour ids outside [0, N) are ignored. */
/* Faces of each site's own list that can never bound a ray's interval inside
 * the site's ball, moved to the end of its range (no reference counterpart):
 * face f is dropped when the site's centre lies strictly on the cell side of
 * f and either f's plane misses the ball by margin * r, or the disc f's plane
 * cuts from the ball lies outside another face g (centre on g's cell side)
 * by margin * r. Cell and ball then intersect exactly as with every face.
 * Writes adjacency_out [E] (and diff_out [E,4] when adjacency_diff, the
 * pf_adj_encode encoding, is given; NULL: faces from the full-precision
 * spheres) with each site's kept faces first in list order, and their number
 * into face_count [N] -- the pf_scene.adjacency_count of the compositing.
 * FP32, eight lanes per site; spheres 16-byte aligned. */
int pf_prune_faces(const float* spheres, int32_t n_sites, const int32_t* adjacency, const int32_t* adjacency_offsets,
                   int64_t n_edges, const uint16_t* adjacency_diff, float margin, int32_t* adjacency_out,
                   uint16_t* diff_out, int32_t* face_count, void* stream);

int pf_power_restrict(const float* spheres, int32_t n_sites, const int32_t* adjacency,
                      const int32_t* adjacency_offsets, int64_t n_edges, uint8_t* keep, int32_t* counts,
                      void* stream);

/* ---- site ordering and the interpenetration regulariser (geometry.py) ----
 * pf_morton_sort: morton_code_kernel / morton_sort (geometry.py:18-85).
 * points [N,3]; order [N] receives the stable ascending order of the 63-bit
 * Morton codes (21 bits per axis over the points' bounding box, extents
 * floored at 1e-6, x-major interleave); codes [N] (or NULL) the unsorted code
 * of every point. The workspace (pf_morton_workspace_size bytes, 256-byte
 * aligned) holds the bounds, keys and the radix sort's storage. NaN
 * coordinates are left out of the bounds and quantise to 0. */
size_t pf_morton_workspace_size(int32_t n_points);
int pf_morton_sort(const float* points, int32_t n_points, void* workspace, size_t workspace_bytes, int64_t* codes,
                   int32_t* order, void* stream);
/* interpenetration_kernel geometry.py:88-117 and its adjoint (:150-189):
 * areas[i] = sum over i's neighbours j with d < r_i + r_j of (r_i + r_j - d)^2,
 * d = |c_i - c_j|. spheres [N,4]; areas [N] overwritten; grad_spheres [N,4]
 * accumulated (zero it first; xyz and r gradients, no direction term when
 * d = 0). Neighbour ids outside [0, N) are skipped. pairing tells the
 * backward what the lists guarantee: 0 nothing; 1 every list ascending without
 * repeats (entries whose reverse exists are then summed without atomics);
 * 2 symmetric (every entry's reverse exists once, as in Cech lists). */
int pf_interpenetration_fwd(const float* spheres, int32_t n_sites, const int32_t* adjacency,
                            const int32_t* adjacency_offsets, int64_t n_edges, float* areas, void* stream);
int pf_interpenetration_bwd(const float* spheres, int32_t n_sites, const int32_t* adjacency,
                            const int32_t* adjacency_offsets, int64_t n_edges, int32_t pairing,
                            const float* grad_areas, float* grad_spheres, void* stream);

/* ---- adjacency encoding (prefetch_adjacency_kernel rasterize.py:580-606) */
/* diff[e] = binary16(c_adj - c, r_adj) for every edge of every site.
 * n_edges is the length of adjacency and diff (adjacency_offsets[n_sites]);
 * NULL adjacency or diff with n_edges > 0 is rejected, and offsets beyond
 * n_edges end a site's edges there. */
int pf_adj_encode(const float* spheres, int32_t n_sites,
                  const int32_t* adjacency, const int32_t* adjacency_offsets,
                  int64_t n_edges, uint16_t* diff, void* stream);

/* ---- compositing (forward_kernel rasterize.py:892-1146,
 *      backward_kernel rasterize.py:1150-1514) --------------------------- */
int pf_raster_fwd(const pf_raster_fwd_args* args, void* stream);
int pf_raster_bwd(const pf_raster_bwd_args* args, void* stream);

/* ---- colour head: degree-3 SH per site, evaluated per texel direction ----
 * No reference counterpart: fills the slot of the reference's per-texel
 * colour pre-pass (SphericalVoronoi._forward / _backward, color_fn.py:188-315)
 * and produces the rasterizer's texel_rgb input.
 * rgb[n,t] = max(0, sh[n,0] + sum_k Y_k(normalize(texel_sites[n,t] - eye)) sh[n,k]).
 * sh is [N,16,3]; texel_sites [N,T,3]; texel_rgb [N,T,3]. */
int pf_sh_fwd(const float* sh, const float* texel_sites, int32_t n_sites, int32_t n_texels, const float* eye,
              float* texel_rgb, void* stream);
/* grad_sh [N,16,3] is overwritten (not accumulated). */
int pf_sh_bwd(const float* sh, const float* texel_sites, const float* grad_texel_rgb, int32_t n_sites,
              int32_t n_texels, const float* eye, float* grad_sh, void* stream);

/* benchmark_kernel rasterize.py:614-734 (Rasterizer.benchmark :1982-2131). */
int pf_raster_rgb(const pf_raster_rgb_args* args, void* stream);

/* visualization_kernel rasterize.py:739-890 (Rasterizer.visualize :2133-2311). */
int pf_raster_vis(const pf_raster_vis_args* args, void* stream);

/* ---- ray tracer (RayTracer.benchmark_kernel raytrace.py:115-241) ------ */
int pf_trace_fwd(const pf_trace_args* args, void* stream);
int pf_trace_bwd(const pf_trace_bwd_args* args, void* stream);

/* Exact face planes for pf_trace_args.face_planes (no reference
 * counterpart: the reference walk forms them from the spheres per visit):
 * face_planes[e] = (n, pm) with n = c_a - c_s and
 * pm = (|c_a|^2 - r_a^2)/2 - (|c_s|^2 - r_s^2)/2 for the face e of site s
 * towards a = adjacency[e], in FP32 as the walk forms them. n_faces = E, the
 * length of scene.adjacency and the float4 capacity of face_planes (16-byte
 * aligned); offsets beyond it (an inconsistent CSR) are not followed. */
int pf_trace_planes(const pf_scene* scene, int64_t n_faces, float* face_planes, void* stream);

/* Longest-walk-first dispatch order for pf_trace_bwd (no reference
 * counterpart: the reference tracer has no backward). tile_order [n_tiles]
 * gets the tiles by descending tile_steps (as pf_trace_fwd writes it), walks
 * of 1023 cells or more sharing the first bucket; the order among equal
 * lengths is unspecified. Waves then start on the long tiles first instead of
 * finishing on them. */
int pf_trace_tile_order(const int32_t* tile_steps, int32_t n_tiles, int32_t* tile_order, void* stream);

/* FP64 recomputation of the walk on sampled rays (no reference counterpart:
 * the precision check of the FP32 walk at full scene size). For each ray
 * (pixel index into cam's image, the ray generated as pf_trace_fwd does) and
 * each cell the FP32 walk visited (cells/n_cells as pf_trace_fwd records
 * them, max_cells per ray), in FP64 from the FP32 scene values: the cell's
 * entry and exit distances (sphere interval cut by the exact power faces of
 * scene.adjacency; NaN where the sphere is missed or the step was not
 * taken), the next cell the FP64 walk would choose (argmin exit-face time,
 * first in list order on ties; 0x7FFFFFFF none, -1 past the walk) and the
 * gap between the smallest and second-smallest exit-face times, with the
 * sum of the FP32 rounding bounds of those two times (the unit roundoff
 * times the magnitudes t = (pm - o.n) / (d.n) is formed from): where the
 * gap is below it, the FP32 walk may order the two faces either way. */
typedef struct pf_trace_check_args {
  pf_camera cam;
  pf_scene scene;            /* spheres, adjacency, adjacency_offsets */
  int32_t n_rays;
  const int32_t* pixels;     /* [R] row-major pixel index i*W + j */
  const int32_t* cells;      /* [R,K] visited cells of those rays */
  const int32_t* n_cells;    /* [R] */
  int32_t max_cells;         /* K */
  double* t_enter;           /* [R,K] */
  double* t_exit;            /* [R,K] */
  int32_t* next_cell;        /* [R,K] */
  double* exit_gap;          /* [R,K] */
  double* exit_tol;          /* [R,K] */
} pf_trace_check_args;
int pf_trace_check_f64(const pf_trace_check_args* args, void* stream);

/* ---- learned colour head (no reference counterpart) -------------------
 * A shared MLP fills the texel-colour slot that the SH head (pf_sh_fwd) and
 * the reference's SphericalVoronoi pre-pass (color_fn.py:54-184) fill: for a
 * row of site s seen along unit direction d,
 *   x = [z_s (16) ; 1, Y_1(d) .. Y_15(d)]   (Y: real SH basis, sh_utils.py:4-49)
 *   rgb = max(0, W3 relu(W2 relu(W1 x + b1) + b2) + b3).
 * W1 [64,32], W2 [64,64], W3 [3,64] and the latents z [N,16] are bf16
 * (row-major, 16-byte aligned); biases fp32. x, h1, h2 are rounded to bf16
 * before each product; products accumulate in fp32. */
typedef struct pf_mlp_weights {
  const uint16_t* w1;
  const float* b1;
  const uint16_t* w2;
  const float* b2;
  const uint16_t* w3;
  const float* b3;
} pf_mlp_weights;

/* Rows: texel mode (pair_sites NULL): row r is texel r of site r / n_texels,
 * d = normalize(texel_sites[r] - eye) (color_fn.py:72). Pair mode: row r is
 * site pair_sites[r] along pair_dirs[r] (training pairs). */
typedef struct pf_mlp_rows {
  int32_t n_rows;
  int32_t n_sites;   /* latents rows; pair rows naming other sites are skipped */
  int32_t n_texels;
  const float* texel_sites; /* [n_rows,3] */
  float eye[3];
  const int32_t* pair_sites; /* [n_rows] or NULL */
  const float* pair_dirs;    /* [n_rows,3] */
  const uint16_t* latents;   /* [N,16] bf16 */
  /* direction encoding e(d), 16 features: PF_MLP_ENC_SH (1, Y_1..Y_15: the
   * SH head's basis) or PF_MLP_ENC_FREQ (1, d, sin(pi d), cos(pi d),
   * sin(2 pi d), cos(2 pi d), per component) */
  int32_t encoding;
  /* texel mode with n_rows = n_sites * n_texels: when set, only the rows of
   * the *site_count sites in site_list (device memory, e.g. from
   * pf_mlp_active_sites or pf_mask_sites) are processed. pf_mlp_fwd writes
   * their rows only; for pf_mlp_bwd the other sites' rows must have a zero
   * colour gradient (they add nothing) and keep the caller's zero latent
   * gradient. NULL: every row. */
  const int32_t* site_list;
  const int32_t* site_count;
  /* pf_mlp_fwd: floats between consecutive output rows, 0 or 3 for rgb
   * [n_rows,3], 4 to write the colour into (rgb, height) records [n_rows,4]
   * (the texel_rgbh layout; heights untouched). */
  int32_t out_stride;
} pf_mlp_rows;
#define PF_MLP_ENC_SH 0
#define PF_MLP_ENC_FREQ 1

/* fp32 gradient buffers; the caller zeroes them first (weights and biases
 * are accumulated into; a texel-mode site whose rows one wave step covers has
 * its latent gradient stored). Bias gradients sum the bf16-rounded output
 * gradients of their layer, as the weight gradients use them. */
typedef struct pf_mlp_grads {
  float* latents; /* [N,16] */
  float* w1;
  float* b1;
  float* w2;
  float* b2;
  float* w3;
  float* b3;
} pf_mlp_grads;

/* The head fused into the compositing forward (pf_raster_fwd_args.mlp_head).
 * Each site that shades a tile gets its T texel colours from the head on the
 * matrix cores inside the compositing kernel, with d = normalize(texel_site -
 * cam.eye) as in texel mode; the results are bit-identical to pf_mlp_fwd's
 * rows. The weights are staged once per workgroup in LDS; the colour
 * channels of scene.texel_rgbh (or texel_rgbh_bf16) are not read, the
 * heights are. The kernel writes (rgb, height) of every texel of every site
 * that shades some tile into texel_rgbh_out [N,T,4] FP32 (16-byte aligned;
 * it may be scene.texel_rgbh itself): the rows pf_raster_bwd then reads as
 * its scene.texel_rgbh (texel_rgbh_bf16 NULL). Rows of other sites are left
 * as they are; the backward reads only rows of sites the forward found
 * shading. Its grad_texel_rgb, fed to pf_mlp_bwd in texel mode, gives the
 * latent and weight gradients. */
typedef struct pf_mlp_head {
  pf_mlp_weights weights;
  const uint16_t* latents;  /* [N,16] bf16, 16-byte aligned */
  int32_t encoding;         /* PF_MLP_ENC_SH or PF_MLP_ENC_FREQ */
  float* texel_rgbh_out;    /* [N,T,4] */
} pf_mlp_head;

/* rgb [n_rows,3]. */
int pf_mlp_fwd(const pf_mlp_weights* w, const pf_mlp_rows* rows, float* rgb, void* stream);

/* The sites whose mask byte is non-zero (and whose mask_and byte is, when
 * mask_and is not NULL), appended to site_list (capacity n_sites) in an
 * unspecified order; *site_count (device, zeroed by the caller) receives
 * their number. No reference counterpart: it turns the binning's visibility
 * mask (pf_bin_emit) and the live mask (pf_live_sites) into the site list of
 * a texel-mode pf_mlp_fwd over the sites a frame's tile lists can hold. */
int pf_mask_sites(const uint8_t* mask, const uint8_t* mask_and, int32_t n_sites, int32_t* site_list,
                  int32_t* site_count, void* stream);
/* Adjoint of pf_mlp_fwd for dL/d rgb = grad_rgb [n_rows,3]; no gradient to
 * directions or texel positions (as color_fn.py:181-184). Steps of 32 rows
 * whose incoming gradients are all zero are skipped (they add nothing). */
int pf_mlp_bwd(const pf_mlp_weights* w, const pf_mlp_rows* rows, const float* grad_rgb, pf_mlp_grads* g,
               void* stream);
/* Appends to site_list [n_sites] each site with a non-zero value among its
 * n_texels (1..8) x 3 colour gradients (grad_rgb [n_sites, n_texels, 3]) and adds
 * their number to *site_count (zeroed by the caller); the list's order is
 * arbitrary. The texel rows the rasterizer's backward leaves without a
 * gradient (sites no ray shades) are most rows of a view. */
int pf_mlp_active_sites(const float* grad_rgb, int32_t n_sites, int32_t n_texels, int32_t* site_list,
                        int32_t* site_count, void* stream);

/* The rasterizer's per-site inputs from the separate parameter arrays, as the
 * concatenations of rasterize.py:1619-1624: spheres [n,4] = (points [n,3],
 * radii [n]), nsigmas [n,4] = (normals [n,3], density [n]), texel_rgbh
 * [n,n_texels,4] = (texel_rgb [n,n_texels,3], texel_height [n,n_texels]).
 * Contiguous FP32 inputs; outputs 16-byte aligned. texel_rgb may be NULL
 * (colours 0: the learned head fused into the compositing supplies them). */
int pf_pack_sites(const float* points, const float* radii, const float* normals, const float* density,
                  const float* texel_rgb, const float* texel_height, int32_t n_sites, int32_t n_texels,
                  float* spheres, float* nsigmas, float* texel_rgbh, void* stream);

/* ---- Spherical-Voronoi texel colour (the reference's colour head,
 *      SphericalVoronoi color_fn.py:27-315) ------------------------------
 * Per texel point p with d = normalize(p - eye): grey 0.5 when
 * d . forward < fov_cos_cutoff, else
 *   w_i = exp(-temp_i |d - axis_i|), rgb = max(0, sum_i w_i val_i / sum_i w_i + 0.5)
 * over sv_dof lobes (color_fn.py:54-101). Lobe arrays are lobe-major as the
 * reference's get_att_sv lays them out (scene.py:290-312): att_sites and
 * att_values [sv_dof, P, 3], att_temps [sv_dof, P]. */
typedef struct pf_sv_args {
  int32_t n_points;         /* P (texels) */
  int32_t sv_dof;           /* lobes */
  const float* points;      /* [P,3] texel sites */
  float eye[3];
  float forward[3];         /* unit normalize(up x right) */
  float fov_cos_cutoff;
  const float* att_sites;   /* [D,P,3] unit lobe axes */
  const float* att_values;  /* [D,P,3] */
  const float* att_temps;   /* [D,P] */
} pf_sv_args;

/* values [P,3], weights_sum [P] (0 where cut off; the backward reads it). */
int pf_sv_fwd(const pf_sv_args* args, float* values, float* weights_sum, void* stream);
/* color_fn.py:105-184: gradients of the lobe axes, values and temperatures,
 * written (not accumulated; zero for cut-off texels and for lobes whose
 * normalised weight is below 1e-3); no gradient to the points. */
int pf_sv_bwd(const pf_sv_args* args, const float* weights_sum, const float* values, const float* grad_values,
              float* grad_att_sites, float* grad_att_values, float* grad_att_temps, void* stream);

/* ---- scene glue (scene.py:239-337): the rasterizer's per-site inputs from
 *      the trainable parameters, forward and adjoint, one launch each -------
 *   radii   = softplus(radii_raw, beta=100), density likewise (torch's
 *             softplus: x where beta x > 20)               scene.py:239-240, 287
 *   normal, tangent, bitangent = columns of R(q / |q|), each normalised
 *                                                            scene.py:242-285
 *   texel_sites[n,k] = points[n] + radii[n] (off[n,k,0] t + off[n,k,1] b)
 *                                                            scene.py:321-329
 *   texel_height[n,k] = height_raw[n,k] radii[n]            scene.py:337 */
typedef struct pf_glue_params {
  const float* points;        /* [N,3] */
  const float* radii;         /* [N]   raw (before softplus) */
  const float* density;       /* [N]   raw */
  const float* quaternions;   /* [N,4] (w, x, y, z), 16-byte aligned */
  const float* texel_offsets; /* [N,T,2] in units of the radius, 8-byte aligned */
  const float* texel_height;  /* [N,T] in units of the radius */
} pf_glue_params;

/* Forward outputs; also the incoming gradients of the adjoint (read only
 * there, any may be NULL = zero). */
typedef struct pf_glue_outputs {
  float* radii;        /* [N] */
  float* density;      /* [N] */
  float* normals;      /* [N,3] */
  float* texel_sites;  /* [N,T,3] */
  float* texel_height; /* [N,T] */
} pf_glue_outputs;

/* Gradients of the parameters, written (not accumulated). `points` holds the
 * texel sites' share only: a caller that also feeds the points to the
 * rasterizer adds that gradient. */
typedef struct pf_glue_grads {
  float* points;        /* [N,3] */
  float* radii;         /* [N] */
  float* density;       /* [N] */
  float* quaternions;   /* [N,4], 16-byte aligned */
  float* texel_offsets; /* [N,T,2], 8-byte aligned */
  float* texel_height;  /* [N,T] */
} pf_glue_grads;

int pf_glue_fwd(const pf_glue_params* params, int32_t n_sites, int32_t n_texels, const pf_glue_outputs* out,
                void* stream);
int pf_glue_bwd(const pf_glue_params* params, int32_t n_sites, int32_t n_texels, const pf_glue_outputs* grad_out,
                const pf_glue_grads* grad_params, void* stream);

/* Replaces non-finite entries with 0 (rasterize.py:1954-1958). */
int pf_scrub_nonfinite(float* data, int64_t n, void* stream);

/* Measurement only, no reference counterpart: the VALU issue peak that the
 * compositing kernels' instruction counts are set against. Launches n_blocks
 * workgroups of 256 threads, each lane running eight independent FP32 FMA
 * chains of `iters` steps; writes one float per thread to `out`
 * [n_blocks*256]. Returns the number of v_fma_f32 wave64 instructions the
 * launch issues (time it on `stream`), or a negative pf_status. */
int64_t pf_probe_valu(int32_t n_blocks, int32_t iters, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* POWERFOAM_AMD_H */
