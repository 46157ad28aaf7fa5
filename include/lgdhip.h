/*
 * lgdhip.h -- C-ABI of the MI355X (gfx950) sweep library behind HipSimulator.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, never
 * throws and never takes ownership of a host pointer.  Every function that
 * can fail returns an int: 0 on success, a negative code otherwise
 * (lgdhip_error_string() turns it into text, lgdhip_last_error() gives the
 * detailed message of the last failure on the calling thread).
 *
 * Device buffers are plain device pointers (from lgdhip_malloc or any other
 * HIP allocator).  Streams and events are opaque handles (hipStream_t and
 * hipEvent_t underneath); a NULL stream is the null stream.
 *
 * Reference interfaces replaced (paths relative to the libgeodecomp tree):
 *   - device allocation / copies / launch loop of
 *     src/libgeodecomp/parallelization/cudasimulator.h:337-340 (cudaMalloc),
 *     :490-497 (getGrid D2H), :515-516 (run H2D), :467-470 (nano-step loop)
 *   - the per-cell sweep kernels kernel2D/kernel3D
 *     (cudasimulator.h:82-155, :160-236) for the user cells
 *       JacobiCellFixedHood  src/testbed/performancetests/main.cpp:1182-1208
 *       ConwayCell           src/examples/gameoflife/main.cpp:25-62
 *       D3Q19 BGK Cell       src/examples/latticeboltzmann/main.cpp:16-237
 *   - CUDASoAGrid's device SoA storage and edge fill
 *     src/libgeodecomp/storage/cudasoagrid.h:16-57, :113-506
 */
#ifndef LGDHIP_H
#define LGDHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* error codes                                                         */
/* ------------------------------------------------------------------ */
#define LGDHIP_OK                 0
#define LGDHIP_ERR_INVALID_ARG   -1   /* bad pointer, size or descriptor     */
#define LGDHIP_ERR_UNSUPPORTED   -2   /* shape/tile the kernel cannot run    */
#define LGDHIP_ERR_WORKSPACE     -3   /* workspace too small                 */
#define LGDHIP_ERR_NO_DEVICE     -4   /* no HIP device visible               */
#define LGDHIP_ERR_HIP         -100   /* HIP runtime error: -100 - hipError  */

typedef void *lgdhip_stream_t;  /* hipStream_t */
typedef void *lgdhip_event_t;   /* hipEvent_t  */

/*
 * Device grid descriptor.  A grid stores `members` planes (struct of
 * arrays, member-major) of sx*sy*sz elements each; the logical cell
 * (x,y,z) of member m lives at element
 *     m*sx*sy*sz + (z+oz)*sx*sy + (y+oy)*sx + (x+ox).
 * The `halo` layers around the logical box hold the topology's edge cell
 * (Cube topology: the constant boundary value), like the padding the
 * reference adds on non-wrapping axes (cudasimulator.h:328-335,
 * soagrid.h:578-584).  For 2D grids nz == 1, oz == 0, sz == 1.
 */
/* Member layouts.  Element (x, y, z) of member m (padded coordinates
 * X = x + ox, Y = y + oy, Z = z + oz) is at element index
 *   m * MS + Z * ZS + Y * sx + X
 * LGDHIP_LAYOUT_MEMBER_PLANES: MS = sx*sy*sz, ZS = sx*sy (one whole padded
 *   plane per member, the member-major SoA of libflatarray);
 * LGDHIP_LAYOUT_Z_SLABS: MS = sx*sy, ZS = members*sx*sy (the members of
 *   one z slice stored together: a sweep over a few z slices touches one
 *   compact address range instead of one range per member, which keeps
 *   HBM streams fewer; the D3Q19 sweep runs ~4% faster).
 * Both layouts take members*sx*sy*sz elements. */
#define LGDHIP_LAYOUT_MEMBER_PLANES 0
#define LGDHIP_LAYOUT_Z_SLABS       1
/* A view: a sub-box of another grid's logical box sharing its memory (see
 * lgdhip_grid_view). */
#define LGDHIP_GRID_VIEW            0x1

typedef struct lgdhip_grid {
    int32_t dim;         /* 2 or 3                                  */
    int32_t halo;        /* padding layers on every side (>= 1)     */
    int32_t elem_bytes;  /* bytes per element of one member         */
    int32_t members;     /* number of SoA member planes             */
    int32_t layout;      /* LGDHIP_LAYOUT_*                         */
    int32_t flags;       /* LGDHIP_GRID_* (0 for grids from _make)  */
    int64_t nx, ny, nz;  /* logical extents                         */
    int64_t ox, oy, oz;  /* element offset of logical (0,0,0)       */
    int64_t sx, sy, sz;  /* allocated extents (row pitch, rows, planes) */
} lgdhip_grid;

/* Tuning knobs picked by the autotuner; NULL selects the defaults. */
typedef struct lgdhip_tile {
    int32_t bx, by, bz;        /* work-group shape (threads)           */
    int32_t steps_per_launch;  /* temporal blocking depth              */
    int32_t variant;           /* kernel variant id (family specific)  */
    int32_t flags;             /* LGDHIP_TILE_* bits                   */
} lgdhip_tile;
/* Work-group -> tile order: XCD_SWIZZLE gives each of the 8 XCDs a
 * contiguous range of tiles (neighbouring tiles share one L2),
 * NO_XCD_SWIZZLE maps them in launch order; neither bit: the family's
 * default (Jacobi temporal blocking: swizzled, LBM: launch order). */
#define LGDHIP_TILE_XCD_SWIZZLE    0x1
#define LGDHIP_TILE_NO_XCD_SWIZZLE 0x2
/* LBM: each x row of tiles on one XCD, rows round-robin over the XCDs */
#define LGDHIP_TILE_XCD_ROWS       0x4

/* Parameters of the D3Q19 lid-driven cavity kernel family. */
typedef struct lgdhip_lbm_params {
    float omega;      /* BGK relaxation rate                     */
    float lid_ux;     /* lid velocity (x component, +y face)     */
    float lid_uy;
    float lid_uz;
} lgdhip_lbm_params;

/* ------------------------------------------------------------------ */
/* runtime                                                             */
/* ------------------------------------------------------------------ */
const char *lgdhip_version(void);
const char *lgdhip_error_string(int code);
const char *lgdhip_last_error(void);
int lgdhip_get_device_count(int *count);
int lgdhip_set_device(int device);
/* The calling thread's current device. */
int lgdhip_get_device(int *device);
int lgdhip_device_name(int device, char *buf, int len);
int lgdhip_device_synchronize(void);

int lgdhip_malloc(void **ptr, size_t bytes);
int lgdhip_free(void *ptr);
int lgdhip_host_malloc(void **ptr, size_t bytes);   /* pinned host memory */
int lgdhip_host_free(void *ptr);
/* host<->device copies return after the copy has completed */
int lgdhip_memcpy_h2d(void *dst, const void *src, size_t bytes, lgdhip_stream_t stream);
int lgdhip_memcpy_d2h(void *dst, const void *src, size_t bytes, lgdhip_stream_t stream);
int lgdhip_memcpy_d2d(void *dst, const void *src, size_t bytes, lgdhip_stream_t stream);
/* `height` rows of `width` bytes, device rows `spitch` bytes apart, host rows
 * `dpitch` apart; synchronous like lgdhip_memcpy_d2h (one transfer for, e.g.,
 * one x-row of every member plane). */
int lgdhip_memcpy2d_d2h(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                        lgdhip_stream_t stream);
/* byte-pattern fill by a kernel (safe inside graph capture) */
int lgdhip_fill_u8(void *dst, uint8_t value, size_t bytes, lgdhip_stream_t stream);
/* streaming copy kernel (roofline / PMC calibration); bytes % 16 == 0 */
int lgdhip_copy_kernel(void *dst, const void *src, size_t bytes, lgdhip_stream_t stream);

int lgdhip_stream_create(lgdhip_stream_t *stream);
int lgdhip_stream_destroy(lgdhip_stream_t stream);
int lgdhip_stream_synchronize(lgdhip_stream_t stream);
int lgdhip_event_create(lgdhip_event_t *event);
int lgdhip_event_destroy(lgdhip_event_t event);
int lgdhip_event_record(lgdhip_event_t event, lgdhip_stream_t stream);
/* Work queued on `stream` after this call waits for `event` (recorded on
 * another stream): orders a copy stream after the launches it reads. */
int lgdhip_stream_wait_event(lgdhip_stream_t stream, lgdhip_event_t event);
int lgdhip_event_synchronize(lgdhip_event_t event);
int lgdhip_event_elapsed_ms(float *ms, lgdhip_event_t start, lgdhip_event_t stop);

/* Graph capture of a launch sequence on `stream` (hipGraph). */
typedef void *lgdhip_graph_t;
int lgdhip_graph_begin(lgdhip_stream_t stream);
int lgdhip_graph_end(lgdhip_stream_t stream, lgdhip_graph_t *graph);
int lgdhip_graph_launch(lgdhip_graph_t graph, lgdhip_stream_t stream);
int lgdhip_graph_destroy(lgdhip_graph_t graph);

/* ------------------------------------------------------------------ */
/* grids (device SoA storage)                                          */
/* ------------------------------------------------------------------ */
/* Fill a descriptor with this library's padding/alignment convention. */
int lgdhip_grid_make(lgdhip_grid *g, int dim, int64_t nx, int64_t ny, int64_t nz,
                     int halo, int elem_bytes, int members);
/* A view of the sub-box [x0, x0+nx) x [y0, y0+ny) x [z0, z0+nz) of g's
 * logical box: same memory, pitches and layout, origin moved to (x0, y0,
 * z0).  Neighbours of the view's cells outside it are g's cells, so a
 * single sweep over a view updates exactly its cells as the full sweep
 * would -- the inner-set / rim split of a region stepper
 * (nesting/cudastepper.h:255-590 updates the rim, exchanges it, then the
 * inner set).  Accepted by lgdhip_grid_load / _store and by the Jacobi
 * single sweeps (variants 1-3; the default picks one); operations that
 * treat the halo as the edge (fills, temporal blocking, Life, LBM) reject
 * views with LGDHIP_ERR_UNSUPPORTED. */
int lgdhip_grid_view(const lgdhip_grid *g, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny,
                     int64_t nz, lgdhip_grid *view);
/* Switch a descriptor made by lgdhip_grid_make to another member layout
 * (LGDHIP_LAYOUT_*); extents and pitches are unchanged. */
int lgdhip_grid_set_layout(lgdhip_grid *g, int layout);
/* Elements per member plane (sx*sy*sz) and total bytes of a grid. */
int64_t lgdhip_grid_plane_elems(const lgdhip_grid *g);
int64_t lgdhip_grid_bytes(const lgdhip_grid *g);
/* Write `edge` (members*elem_bytes bytes, member-major) into every halo
 * element of every member plane.  Replaces the padding fill of
 * cudasimulator.h:508-513 and cudasoagrid.h:16-57. */
int lgdhip_grid_fill_edge(void *grid, const lgdhip_grid *g, const void *edge,
                          lgdhip_stream_t stream);
/* Refresh every halo layer for a (partly) periodic topology: a halo cell
 * takes the interior value at its coordinate wrapped along the axes in
 * wrap_mask (1 x, 2 y, 4 z) or, if it is still outside the box along an
 * unwrapped axis, `edge` (members*elem_bytes bytes) -- the neighbourhood
 * rule of the Torus topologies (geometry/topologies.h normalize /
 * isOutOfBounds).  Call before each single sweep of a periodic grid. */
int lgdhip_grid_fill_halo(void *grid, const lgdhip_grid *g, const void *edge, int wrap_mask,
                          lgdhip_stream_t stream);
/* ------------------------------------------------------------------ */
/* cell grids (device array-of-structs storage of whole user cells)    */
/* ------------------------------------------------------------------ */
/* The storage of the generic kernel family, which runs any user cell whose
 * update() is compiled by hipcc.  Cells of cell_bytes bytes, x fastest,
 * padded by `halo` cells on both sides of each of the grid's dim axes:
 *   px = nx + 2h,  py = dim >= 2 ? ny + 2h : 1,  pz = dim == 3 ? nz + 2h : 1,
 *   cell (x, y, z) at index ((z + hz) * py + (y + hy)) * px + (x + h),
 *   hy = dim >= 2 ? h : 0, hz = dim == 3 ? h : 0.
 * Replaces the padded AoS device buffers of CUDASimulator
 * (cudasimulator.h:314-429, :502-534). */
typedef struct lgdhip_cell_grid {
    int32_t dim, halo;
    int64_t cell_bytes;
    int64_t nx, ny, nz;
    int64_t px, py, pz;
} lgdhip_cell_grid;
int lgdhip_cell_grid_make(lgdhip_cell_grid *g, int dim, int64_t nx, int64_t ny, int64_t nz, int halo,
                          int64_t cell_bytes);
int64_t lgdhip_cell_grid_bytes(const lgdhip_cell_grid *g);
/* Write every halo cell: the interior cell at its coordinate wrapped along
 * the axes in wrap_mask (1 x, 2 y, 4 z), or, where the coordinate is still
 * outside an unwrapped axis, the cell at `edge` (DEVICE memory, cell_bytes)
 * -- the neighbourhood rule of the reference's topologies
 * (geometry/topologies.h:185-234, replacing the per-thread wrap offsets of
 * cudasimulator.h:52-64).  wrap_mask 0 writes the edge cell everywhere. */
int lgdhip_cell_grid_fill_halo(void *cells, const lgdhip_cell_grid *g, const void *edge, int wrap_mask,
                               lgdhip_stream_t stream);
/* Copy the logical box between a dense host array (x fastest, nx*ny*nz
 * cells) and a device cell grid; both return after the copy completed. */
int lgdhip_cell_grid_load(void *cells, const lgdhip_cell_grid *g, const void *host, lgdhip_stream_t stream);
int lgdhip_cell_grid_store(const void *cells, const lgdhip_cell_grid *g, void *host, lgdhip_stream_t stream);

/* Copy the logical box between a dense host array (member-major, x
 * fastest, nx*ny*nz elements per member) and a device grid. */
int lgdhip_grid_load(void *grid, const lgdhip_grid *g, const void *host,
                     lgdhip_stream_t stream);
int lgdhip_grid_store(const void *grid, const lgdhip_grid *g, void *host,
                      lgdhip_stream_t stream);

/* ------------------------------------------------------------------ */
/* Jacobi3D, FP64, 7-point, Cube topology                              */
/*   T' = (T[z-1]+T[y-1]+T[x-1]+T+T[x+1]+T[y+1]+T[z+1]) * (1.0/7.0)     */
/*   JacobiCellFixedHood::update, main.cpp:1194-1205                   */
/* ------------------------------------------------------------------ */
/* Runs `steps` sweeps ping-ponging between a and b (both carry the edge
 * value in their halo).  *result_in_b is set to 1 when the final state is
 * in b, 0 when it is in a (tile->variant 4-10 run K = 2-4 sweeps per launch
 * and swap once per launch, so this is not steps % 2 there; 11 and 12 run
 * the same tile pipelines for one sweep per launch). */
int lgdhip_jacobi3d_f64_run(double *a, double *b, const lgdhip_grid *g, int steps,
                            const lgdhip_tile *tile, int *result_in_b,
                            lgdhip_stream_t stream);

/* The launch plan lgdhip_jacobi3d_f64_run takes for this grid and tile
 * (tile may be NULL): *variant gets the kernel variant (the default resolves
 * to 6 or 10 by the launch cost model), *zchunk the z planes per work-group.
 * No device work.  Build-side query for tools (benchmarks, tuners); the
 * reference's CUDASimulator has no counterpart (its launch shape is the
 * block size it is given, cudasimulator.h:314-429). */
int lgdhip_jacobi3d_f64_plan(const lgdhip_grid *g, const lgdhip_tile *tile, int *variant, int *zchunk);

/* One launch of a temporal-blocking variant (tile->variant 4-12, or the
 * default plan's) writing only the planes [z0, z1) of dst: K sweeps of src
 * for those planes (the rest of dst is not touched), *sweeps = K.  The
 * building block of a slab stepper that updates its rim planes first, hands
 * them to the neighbours' ghost zones, and updates the inner planes while
 * the copies run (the inner-set / rim split of nesting/cudastepper.h:336-380,
 * 516-575).  Whole grids only. */
int lgdhip_jacobi3d_f64_run_zrange(const double *src, double *dst, const lgdhip_grid *g, int z0, int z1,
                                   const lgdhip_tile *tile, int *sweeps, lgdhip_stream_t stream);

/* The same sweeps with periodic axes: wrap_mask bits 1, 2, 4 wrap x, y, z
 * (Topologies::Torus and the selectively wrapped topologies,
 * geometry/topologies.h:225-247,267-280: a neighbour past a wrapped face is
 * the cell on the opposite face; past an unwrapped face it is `edge`).  The
 * temporal-blocking variants take the wrapped neighbours straight from the
 * interior; steps % K are single sweeps after a periodic halo refresh
 * (lgdhip_grid_fill_halo).  Both buffers' halos are rewritten.  A wrapped z
 * axis shorter than 8 cells runs single sweeps throughout.  Results equal
 * `steps` single sweeps of the periodic stencil bit for bit; *result_in_b
 * as for lgdhip_jacobi3d_f64_run.  Views are refused. */
int lgdhip_jacobi3d_f64_run_periodic(double *a, double *b, const lgdhip_grid *g, int steps, int wrap_mask,
                                     double edge, const lgdhip_tile *tile, int *result_in_b,
                                     lgdhip_stream_t stream);

/* ------------------------------------------------------------------ */
/* Game of Life, uint8 cells (0/1), Moore<2,1>, Cube topology          */
/*   ConwayCell::update, gameoflife/main.cpp:36-59                     */
/* ------------------------------------------------------------------ */
/* Byte-per-cell sweep, one generation per launch; ping-pongs a/b. */
int lgdhip_life2d_u8_run(uint8_t *a, uint8_t *b, const lgdhip_grid *g, int steps,
                         int *result_in_b, lgdhip_stream_t stream);
/* Bit-packed state: ceil(nx/32) 32-bit words per row, ny rows, bit i of
 * word w is cell x = 32*w + i. */
int64_t lgdhip_life2d_packed_words(const lgdhip_grid *g);
int lgdhip_life2d_pack(const uint8_t *grid, const lgdhip_grid *g, uint32_t *packed,
                       lgdhip_stream_t stream);
int lgdhip_life2d_unpack(const uint32_t *packed, const lgdhip_grid *g, uint8_t *grid,
                         lgdhip_stream_t stream);
/* `steps` generations on the packed state (register temporal blocking,
 * tile->steps_per_launch generations per launch); edge_alive is the
 * constant state of every cell outside the grid. */
int lgdhip_life2d_packed_run(uint32_t *pa, uint32_t *pb, const lgdhip_grid *g,
                             int edge_alive, int steps, const lgdhip_tile *tile,
                             int *result_in_b, lgdhip_stream_t stream);
/* The same generations with periodic axes: wrap_mask bit 1 wraps x (needs
 * nx % 32 == 0: the packed columns wrap word by word), bit 2 wraps y
 * (Topologies::Torus and the cylinders, geometry/topologies.h:225-247,
 * 267-280); cells past an unwrapped face hold the edge state.  The blocked
 * kernels load the periodic images of rows and word columns directly. */
int lgdhip_life2d_packed_run_periodic(uint32_t *pa, uint32_t *pb, const lgdhip_grid *g, int edge_alive, int steps,
                                      int wrap_mask, const lgdhip_tile *tile, int *result_in_b,
                                      lgdhip_stream_t stream);
/* uint8 grid in, uint8 grid out (in place): pack -> packed_run -> unpack.
 * `workspace` must hold lgdhip_life2d_workspace_bytes(g) bytes. */
size_t lgdhip_life2d_workspace_bytes(const lgdhip_grid *g);
int lgdhip_life2d_run(uint8_t *grid, const lgdhip_grid *g, int edge_alive, int steps,
                      void *workspace, size_t workspace_bytes,
                      const lgdhip_tile *tile, lgdhip_stream_t stream);

/* ------------------------------------------------------------------ */
/* D3Q19 BGK lid-driven cavity, FP32 populations (19 SoA planes)       */
/*   fused pull stream+collide after latticeboltzmann/main.cpp:92-229  */
/* ------------------------------------------------------------------ */
/* Populations are stored shifted, g_i = f_i - w_i (w = 1/3, 1/18, 1/36 for
 * rest, axis and diagonal links), 19 member planes in the order
 * C N E W S T B NW SW NE SE TW BW TE BE TN BN TS BS.  `flags` is a uint8
 * grid made with lgdhip_grid_make(3, nx, ny, nz, halo, 1, 1) holding the
 * LGDHIP_LBM_* wall bits of every cell.  Ping-pongs a/b like the other
 * families.  tile (optional): by = rows per 64-wide work-group (1, 2, 4, 8),
 * or bx = 256 for 256 x 1 work-groups; variant 1 = nontemporal population
 * loads; flags (LGDHIP_TILE_*). */
#define LGDHIP_LBM_WALL_XLO 0x01
#define LGDHIP_LBM_WALL_XHI 0x02
#define LGDHIP_LBM_WALL_YLO 0x04
#define LGDHIP_LBM_WALL_YHI 0x08
#define LGDHIP_LBM_WALL_ZLO 0x10
#define LGDHIP_LBM_WALL_ZHI 0x20
#define LGDHIP_LBM_LID_YHI  0x40
int lgdhip_lbm_d3q19_f32_run(float *a, float *b, const uint8_t *flags,
                             const lgdhip_grid *g, const lgdhip_lbm_params *p,
                             int steps, const lgdhip_tile *tile, int *result_in_b,
                             lgdhip_stream_t stream);
/* Cavity wall flags for the whole box (lid on the +y face). */
int lgdhip_lbm_cavity_flags(uint8_t *flags, const lgdhip_grid *fg, lgdhip_stream_t stream);
/* rho = 1 + sum g and the momentum j = sum e g (the incompressible model's
 * velocity) as dense nx*ny*nz float device arrays. */
int lgdhip_lbm_d3q19_f32_moments(const float *f, const lgdhip_grid *g, float *rho,
                                 float *ux, float *uy, float *uz, lgdhip_stream_t stream);

/* ------------------------------------------------------------------ */
/* Batched ensemble of 3D Poisson problems, Jacobi 7-point             */
/*   6 u_c - sum_nb u_nb = f_c, constant Dirichlet value per face.     */
/*   No reference counterpart (the sweep is the Jacobi cell of         */
/*   src/testbed/performancetests/main.cpp:1194-1205 plus a source).   */
/* ------------------------------------------------------------------ */
/* Problem k occupies a padded (n+2)^3 block (x fastest) starting at
 * element `offset` of the pools: u (double; its halo holds the faces) and
 * f (float; halo unused). */
typedef struct lgdhip_poisson_problem {
    int32_t n;          /* interior cells per edge                      */
    int32_t reserved;
    int64_t offset;     /* element offset of the padded block           */
    double face[6];     /* Dirichlet values at x=-1, x=n, y=-1, y=n, z=-1, z=n */
} lgdhip_poisson_problem;

typedef struct lgdhip_poisson_options {
    double rel_tol;              /* stop at ||r||_2 <= rel_tol * reference        */
    int32_t chunk;               /* FP32 sweeps between convergence checks (even) */
    int32_t max_sweeps;          /* per problem                                    */
    double inner_reduction;      /* FP64 refresh once the correction's residual
                                    fell by this factor (0: 1e-2)                 */
    int32_t use_graph;           /* replay chunks as a captured hipGraph           */
    int32_t sweep_variant;       /* 0 default (6 for active sets of at least
                                    2048 of its tiles, else 3), 1 plain sweeps,
                                    2 / 3 temporal blocking with 2 / 4 sweeps per
                                    launch, 4 / 5 4 sweeps with two rows per
                                    thread (32- / 16-row tiles), 6 with three
                                    rows (48-row tiles)                           */
    const double *reference_norms; /* host array, one per problem, or NULL: the
                                      residual norm of the initial guess          */
    int32_t smooth_sweeps;       /* weighted-Jacobi sweeps before the plain ones
                                    (0: none; rounded up to even)                 */
    float smooth_omega;          /* their weight in (0, 1], e.g. 2/3              */
    int32_t final_fp64_sweeps;   /* plain FP64 Jacobi sweeps on u after every
                                    problem met the tolerance (0: none; rounded up
                                    to even); the FP64 residual is recomputed and
                                    reported after them                           */
} lgdhip_poisson_options;

typedef struct lgdhip_poisson_result {
    int64_t sweeps;              /* Jacobi sweeps performed                        */
    int32_t outer_iterations;    /* FP64 refresh steps                             */
    int32_t converged;
    double reference_norm;
    double residual_norm;        /* final FP64 ||f + sum_nb u - 6u||_2             */
} lgdhip_poisson_result;

/* Assigns consecutive, 256-byte aligned offsets; returns the pool length in
 * elements (allocate u with 8 and f with 4 bytes per element). */
int64_t lgdhip_poisson_layout(lgdhip_poisson_problem *probs, int count);
size_t lgdhip_poisson_workspace_bytes(const lgdhip_poisson_problem *probs, int count);
/* Writes every problem's face values into the halo of u. */
int lgdhip_poisson_set_faces(double *u, const lgdhip_poisson_problem *probs, int count, void *workspace,
                             size_t workspace_bytes, lgdhip_stream_t stream);
/* FP64 residual norms ||f + sum_nb u - 6u||_2 into the host array norms. */
int lgdhip_poisson_residual_norms(const double *u, const float *f, const lgdhip_poisson_problem *probs,
                                  int count, void *workspace, size_t workspace_bytes, double *norms,
                                  lgdhip_stream_t stream);
/* The hot sweep alone: `sweeps` FP32 Jacobi sweeps on A e = r for every
 * problem (e pools with zero halos, ping-pong ea/eb).  tile->variant as
 * lgdhip_poisson_options::sweep_variant (NULL: default); blocked launches
 * swap the buffers once per 2 or 4 sweeps, so *result_in_b is not sweeps % 2
 * there. */
int lgdhip_poisson_sweeps_f32(float *ea, float *eb, const float *r, const lgdhip_poisson_problem *probs, int count,
                              int sweeps, const lgdhip_tile *tile, void *workspace, size_t workspace_bytes,
                              int *result_in_b, lgdhip_stream_t stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  lgdhip_poisson_sweeps_f32 without the workspace: the
 * problem descriptors and tile lists go into a static 8 MiB device array
 * per GPU (the whole 512-problem ensemble of n in 32..192 needs ~5 MiB);
 * a problem set that needs more fails with LGDHIP_ERR_WORKSPACE.  Calls no
 * allocating routine.  Calls that overlap on one GPU share that array and
 * race; serialise them (one stream, or one host thread per GPU). */
int lgdhip_poisson_sweeps_f32_noalloc(float *ea, float *eb, const float *r, const lgdhip_poisson_problem *probs,
                                      int count, int sweeps, const lgdhip_tile *tile, int *result_in_b,
                                      lgdhip_stream_t stream);

/* `sweeps` plain FP64 Jacobi sweeps on u (ping-pong ua/ub, both with the
 * faces in their halos, e.g. from lgdhip_poisson_set_faces), each cell
 * (u[z-1] + u[y-1] + u[x-1] + u[x+1] + u[y+1] + u[z+1] + f) * (1/6) summed in
 * that order; the final polishing sweeps of lgdhip_poisson_batch_solve.
 * *result_in_b = sweeps % 2.  No reference counterpart (the ensemble
 * solver is new with this library). */
int lgdhip_poisson_sweeps_f64(double *ua, double *ub, const float *f, const lgdhip_poisson_problem *probs,
                              int count, int sweeps, void *workspace, size_t workspace_bytes, int *result_in_b,
                              lgdhip_stream_t stream);

/* The blocked sweep variant a launch over these problems uses for
 * `requested` (0: the default, resolved by launch size; 1-6 are kept).  No
 * device work; for tools.  No counterpart in the reference (the ensemble
 * solver is new with this library). */
int lgdhip_poisson_sweep_plan(const lgdhip_poisson_problem *probs, int count, int requested, int *variant);

/* Weighted (damped) sweeps e' = e + omega (J(e) - e), omega in (0, 1]: as
 * lgdhip_poisson_sweeps_f32 (same variants and buffer convention), each cell
 * computing fma(omega, fma(sum, 1/6, -e), e) with single roundings.  The
 * smoothing phase of lgdhip_poisson_batch_solve runs these. */
int lgdhip_poisson_damped_sweeps_f32(float *ea, float *eb, const float *r, const lgdhip_poisson_problem *probs,
                                     int count, int sweeps, float omega, const lgdhip_tile *tile, void *workspace,
                                     size_t workspace_bytes, int *result_in_b, lgdhip_stream_t stream);
/* Solves every problem from the guess in u (interior) to opt->rel_tol:
 * FP64 residual and update, FP32 Jacobi sweeps on the correction, all
 * active problems swept by one launch per sweep. */
int lgdhip_poisson_batch_solve(double *u, const float *f, const lgdhip_poisson_problem *probs, int count,
                               const lgdhip_poisson_options *opt, void *workspace, size_t workspace_bytes,
                               lgdhip_poisson_result *results, lgdhip_stream_t stream);

/* ------------------------------------------------------------------ */
/* 3D U-Net building blocks for the learned initial guess             */
/*   (no reference counterpart): BF16 operands on the MFMA cores,     */
/*   FP32 accumulation.                                                */
/* ------------------------------------------------------------------ */
/* Activations are channels-last BF16 cubes: element (z,y,x,c) of an n^3
 * tensor with channel stride ld and channel offset off is at
 * ((z*n+y)*n+x)*ld + off + c.  Weights are BF16 [cout][Kp], k = tap*cin + ci,
 * tap = (dz+1)*9 + (dy+1)*3 + (dx+1) (3x3x3) or 0 (1x1x1), zero padded to
 * Kp = lgdhip_conv3d_kpad(cin, ksize).  Zero padding at the faces. */
int lgdhip_conv3d_kpad(int cin, int ksize);
/* y[v][co] (+)= act(sum_k x[v+tap][ci] w[co][k] + bias[co]); BF16 output to y
 * (if non-NULL) and/or FP32 output to y32 (channel stride ld32).  relu != 0
 * clamps at 0; accumulate != 0 adds the BF16 value already in y.  cout <= 192;
 * when cin % 8 == 0, x must be 16-byte aligned and ldx, offx multiples of 8. */
int lgdhip_conv3d_bf16_forward(const void *x, int n, int ldx, int offx, int cin, int ksize, const void *w,
                               const float *bias, int cout, void *y, int ldy, int offy, float *y32, int ld32,
                               int relu, int accumulate, lgdhip_stream_t stream);
/* The same convolution with a workspace for the small volumes: 3x3x3 layers
 * with cin % 8 == 0 below 64^3 split their k range over work-groups and add
 * the FP32 partials in a fixed order (deterministic; the k order differs
 * from the unsplit launch, so results agree to FP32 rounding, not bit for
 * bit).  workspace_bytes >= lgdhip_conv3d_forward_workspace_bytes(n, cin,
 * ksize, cout), which is 0 for layers that do not split; a NULL workspace
 * runs the unsplit kernels, as lgdhip_conv3d_bf16_forward does. */
size_t lgdhip_conv3d_forward_workspace_bytes(int n, int cin, int ksize, int cout);
int lgdhip_conv3d_bf16_forward_ws(const void *x, int n, int ldx, int offx, int cin, int ksize, const void *w,
                                  const float *bias, int cout, void *y, int ldy, int offy, float *y32, int ld32,
                                  int relu, int accumulate, void *workspace, size_t workspace_bytes,
                                  lgdhip_stream_t stream);
/* weights of the input-gradient convolution: wt[ci][k'] (Kp' = kpad(cout)),
 * flipped taps, so dx = conv(dy, wt) with cin' = cout, cout' = cin */
int lgdhip_conv3d_dgrad_weights(const void *w, int cin, int cout, int ksize, void *wt, lgdhip_stream_t stream);
/* dw[co][k] (+)= sum_v dy[v][co] x[v+tap(k)][ci(k)] (FP32, fixed-order
 * reduction over voxel chunks of chunk_vox voxels, a multiple of 32; <= 0:
 * the library's choice for the shape, the largest power of two <= 4096 and
 * >= 128 that gives 2048 waves -- pass the same value to the workspace
 * query), db[co] = sum_v dy[v][co] (if non-NULL) */
size_t lgdhip_conv3d_wgrad_workspace_bytes(int n, int cin, int ksize, int cout, int chunk_vox);
int lgdhip_conv3d_bf16_wgrad(const void *x, int n, int ldx, int offx, int cin, int ksize, const void *dy, int ldy,
                             int offy, int cout, float *dw, float *db, int accumulate, int chunk_vox,
                             void *workspace, size_t workspace_bytes, lgdhip_stream_t stream);
/* g *= (y > 0), in place */
int lgdhip_relu_backward_bf16(void *g, int ldg, int offg, const void *y, int ldy, int offy, int c, int64_t nvox,
                              lgdhip_stream_t stream);
/* 2x2x2 average pool n^3 -> (n/2)^3, its gradient (accumulate adds) */
int lgdhip_avgpool2_bf16(const void *x, int n, int ldx, int offx, void *y, int ldy, int offy, int c,
                         lgdhip_stream_t stream);
int lgdhip_avgpool2_backward_bf16(const void *dy, int n, int ldy, int offy, void *dx, int ldx, int offx, int c,
                                  int accumulate, lgdhip_stream_t stream);
/* nearest 2x upsample m^3 -> (2m)^3, its gradient (sum of the 8 children) */
int lgdhip_upsample2_bf16(const void *x, int m, int ldx, int offx, void *y, int ldy, int offy, int c,
                          lgdhip_stream_t stream);
int lgdhip_upsample2_backward_bf16(const void *dy, int m, int ldy, int offy, void *dx, int ldx, int offx, int c,
                                   int accumulate, lgdhip_stream_t stream);
/* nearest 2x upsample of x (m^3, cu channels) and a skip tensor ((2m)^3, cs
 * channels) written as whole concatenation rows y[v][offy .. offy+cu+cs)
 * (the upsampled channels first); no reference counterpart: the U-Net of
 * the learned initial guess */
int lgdhip_upsample2_concat_bf16(const void *x, int m, int ldx, int offx, int cu, const void *skip, int lds,
                                 int offs, int cs, void *y, int ldy, int offy, lgdhip_stream_t stream);
/* FP32 field -> one BF16 channel of a channels-last tensor */
int lgdhip_pack_bf16(const float *src, int64_t nvox, void *y, int ldy, int offy, lgdhip_stream_t stream);
/* MSE: grad[v*ldg] = 2 (pred - target) / nvox (BF16), partial[b] = per-block
 * sums of squared errors (lgdhip_mse_partials(nvox) entries) */
int64_t lgdhip_mse_partials(int64_t nvox);
int lgdhip_mse_loss(const float *pred, const float *target, int64_t nvox, void *grad, int ldg, float *partial,
                    lgdhip_stream_t stream);
/* Adam on FP32 master weights (kp > 0: [rows][kp] with the columns
 * >= cols_used kept zero); writes BF16 copies to w_bf16 if non-NULL */
int lgdhip_adam_step(float *w, float *m, float *v, const float *g, int64_t count, int cols_used, int kp, float lr,
                     float beta1, float beta2, float eps, int t, void *w_bf16, lgdhip_stream_t stream);

/* ------------------------------------------------------------------ */
/* Learned initial guess: Poisson pools <-> coarse grid <-> U-Net      */
/*   tensors (no reference counterpart).  f and u are the pool base    */
/*   pointers; prob (host) gives n, the block offset and the faces.    */
/*   symmetry in [0,48) reads the problem through a cube symmetry:     */
/*   axis permutation (symmetry % 6) and reflections (bits of          */
/*   symmetry / 6).  A base field is dense n^3 FP32 (x fastest).       */
/* ------------------------------------------------------------------ */
/* x[v][0] = bf16(f/s), x[v][1] = bf16(b/s) with b the face field (sum of the
 * Dirichlet values of the faces cell v touches), x[v][2] = bf16(base/s) when
 * base != NULL, remaining channels up to ldx zero; inv_scale = 1/s */
int lgdhip_poisson_unet_input(const float *f, const lgdhip_poisson_problem *prob, double inv_scale, int symmetry,
                              const float *base, void *x, int ldx, lgdhip_stream_t stream);
/* t[v] = (float)((u_interior[v] - base[v]) / s) (base = 0 when NULL): the
 * training target */
int lgdhip_poisson_unet_target(const double *u, const lgdhip_poisson_problem *prob, double inv_scale, int symmetry,
                               const float *base, float *t, lgdhip_stream_t stream);
/* u_interior[v] = pred[v] * s + base[v] (either may be NULL, not both): the
 * initial guess; the halo is left alone */
int lgdhip_poisson_unet_guess(const float *pred, const lgdhip_poisson_problem *prob, double scale, const float *base,
                              double *u, lgdhip_stream_t stream);
/* Coarse grid of a problem: m^3 cells of spacing H = (n+1)/(m+1) at fine
 * positions (j+1)H - 1, its halo on the fine halo planes (same faces); the
 * coarse equation is 6U - sum U_nb = H^2 f_c. */
double lgdhip_poisson_coarse_spacing(int n, int m);
/* fc (coarse pool, block coarse->offset) = H^2 x hat-weighted average of f
 * around each coarse cell (weights max(0, 1 - |i - X|/H) per axis) */
int lgdhip_poisson_coarse_restrict(const float *f, const lgdhip_poisson_problem *prob, float *fc,
                                   const lgdhip_poisson_problem *coarse, lgdhip_stream_t stream);
/* out (dense n^3 FP32) = trilinear interpolation of the coarse solution uc
 * (coarse pool, halo = faces; edge/corner halo cells read as the mean of the
 * faces meeting there) at the fine cells */
int lgdhip_poisson_coarse_prolong(const double *uc, const lgdhip_poisson_problem *coarse, int n, float *out,
                                  lgdhip_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* LGDHIP_H */
