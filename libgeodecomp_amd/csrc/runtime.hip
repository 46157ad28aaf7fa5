// Runtime part of the lgdhip C-ABI: errors, memory, streams, events, graphs
// and the device SoA grid helpers (edge fill, host<->device box copies).
//
// Replaces the inline CUDA runtime calls of the reference simulator
// (src/libgeodecomp/parallelization/cudasimulator.h:337-340 cudaMalloc,
// :515-516 upload, :490-497 download) and the padding fill of CUDASoAGrid
// (src/libgeodecomp/storage/cudasoagrid.h:16-57).
#include "lgdhip_internal.h"

#include <cstring>
#include <vector>

namespace lgdhip {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    return set_error(hip_code(e), m);
}

int check_grid(const lgdhip_grid *g, int dim, int elem_bytes, int members, bool allow_view) {
    if (!g) return set_error(LGDHIP_ERR_INVALID_ARG, "grid descriptor is NULL");
    if (g->flags & ~LGDHIP_GRID_VIEW) return set_error(LGDHIP_ERR_INVALID_ARG, "unknown grid flags");
    if ((g->flags & LGDHIP_GRID_VIEW) && !allow_view)
        return set_error(LGDHIP_ERR_UNSUPPORTED, "this operation needs a whole grid, not a view");
    if (dim && g->dim != dim) return set_error(LGDHIP_ERR_INVALID_ARG, "grid has the wrong dimensionality");
    if (elem_bytes && g->elem_bytes != elem_bytes)
        return set_error(LGDHIP_ERR_INVALID_ARG, "grid has the wrong element size");
    if (members && g->members != members)
        return set_error(LGDHIP_ERR_INVALID_ARG, "grid has the wrong number of members");
    if (g->dim != 2 && g->dim != 3) return set_error(LGDHIP_ERR_INVALID_ARG, "grid dim must be 2 or 3");
    if (g->layout != LGDHIP_LAYOUT_MEMBER_PLANES && g->layout != LGDHIP_LAYOUT_Z_SLABS)
        return set_error(LGDHIP_ERR_INVALID_ARG, "unknown grid layout");
    if (g->nx < 1 || g->ny < 1 || g->nz < 1) return set_error(LGDHIP_ERR_INVALID_ARG, "empty grid");
    if (g->halo < 1) return set_error(LGDHIP_ERR_INVALID_ARG, "grid halo must be >= 1");
    const int64_t h = g->halo;
    if (g->ox < h || g->oy < h) return set_error(LGDHIP_ERR_INVALID_ARG, "grid offset smaller than halo");
    if (g->sx < g->ox + g->nx + h || g->sy < g->oy + g->ny + h)
        return set_error(LGDHIP_ERR_INVALID_ARG, "grid pitch too small for the halo");
    if (g->dim == 3) {
        if (g->oz < h || g->sz < g->oz + g->nz + h)
            return set_error(LGDHIP_ERR_INVALID_ARG, "grid z extent too small for the halo");
    } else {
        if (g->nz != 1 || g->oz != 0 || g->sz != 1)
            return set_error(LGDHIP_ERR_INVALID_ARG, "2D grid must have nz == sz == 1 and oz == 0");
    }
    // every index the kernels form must fit 63 bits and every extent 31 bits
    if (g->nx > (1 << 30) || g->ny > (1 << 30) || g->nz > (1 << 30))
        return set_error(LGDHIP_ERR_UNSUPPORTED, "grid extent above 2^30");
    return 0;
}

} // namespace lgdhip

using namespace lgdhip;

// ---------------------------------------------------------------------------
// kernels used by the runtime
// ---------------------------------------------------------------------------
__global__ void lgd_fill_u8_kernel(uint8_t *__restrict__ dst, uint8_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

__global__ void lgd_fill_u32_kernel(uint32_t *__restrict__ dst, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

// one 16-byte element per thread, no grid-stride loop: the fastest plain
// streaming copy measured on MI355X (~6.1 TB/s read+write for 2-3 GiB)
__global__ __launch_bounds__(256) void lgd_copy16_kernel(uint4 *__restrict__ dst,
                                                         const uint4 *__restrict__ src, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Writes the edge value into every halo element of one member (z slices
// `zs` elements apart).  One thread per allocated element of the member;
// interior elements are left alone.
template <typename T>
__global__ void lgd_fill_edge_kernel(T *__restrict__ plane, T edge, int64_t sx, int64_t sy, int64_t sz,
                                     int64_t zs, int64_t ox, int64_t oy, int64_t oz, int64_t nx, int64_t ny,
                                     int64_t nz) {
    const int64_t n = sx * sy * sz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = i % sx;
        const int64_t y = (i / sx) % sy;
        const int64_t z = i / (sx * sy);
        const bool inside = x >= ox && x < ox + nx && y >= oy && y < oy + ny && z >= oz && z < oz + nz;
        if (!inside) plane[z * zs + y * sx + x] = edge;
    }
}

// Refreshes the halo layers of one member plane for a (partly) periodic
// topology: a halo cell takes the interior value at its coordinate wrapped
// along the axes set in `wrap` (bit 0 x, 1 y, 2 z) or, when it is still
// outside the box along an unwrapped axis, the edge value -- the rule of
// Topology::normalize + isOutOfBounds.  Reads only interior cells, writes
// only halo cells.
template <typename T>
__global__ void lgd_fill_halo_kernel(T *__restrict__ plane, T edge, int64_t sx, int64_t zs, int64_t ox, int64_t oy,
                                     int64_t oz, int64_t nx, int64_t ny, int64_t nz, int h, int dim, int wrap) {
    const int64_t ex = nx + 2 * h, ey = ny + 2 * h, ez = dim == 3 ? nz + 2 * h : 1;
    const int64_t total = ex * ey * ez;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = i % ex - h, y = (i / ex) % ey - h, z = dim == 3 ? i / (ex * ey) - h : 0;
        if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) continue;
        int64_t wx = x, wy = y, wz = z;
        if (wrap & 1) wx = ((x % nx) + nx) % nx;
        if (wrap & 2) wy = ((y % ny) + ny) % ny;
        if ((wrap & 4) && dim == 3) wz = ((z % nz) + nz) % nz;
        const bool inside = wx >= 0 && wx < nx && wy >= 0 && wy < ny && wz >= 0 && wz < nz;
        const T v = inside ? plane[(wz + oz) * zs + (wy + oy) * sx + (wx + ox)] : edge;
        plane[(z + oz) * zs + (y + oy) * sx + (x + ox)] = v;
    }
}

static unsigned fill_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return (unsigned)b;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *lgdhip_version(void) { return "lgdhip 0.1 (gfx950)"; }

const char *lgdhip_error_string(int code) {
    switch (code) {
    case LGDHIP_OK: return "success";
    case LGDHIP_ERR_INVALID_ARG: return "invalid argument";
    case LGDHIP_ERR_UNSUPPORTED: return "unsupported shape or tile";
    case LGDHIP_ERR_WORKSPACE: return "workspace too small";
    case LGDHIP_ERR_NO_DEVICE: return "no HIP device";
    default: break;
    }
    if (code <= LGDHIP_ERR_HIP) return hipGetErrorString((hipError_t)(LGDHIP_ERR_HIP - code));
    return "unknown error";
}

const char *lgdhip_last_error(void) { return g_last_error.c_str(); }

int lgdhip_get_device_count(int *count) {
    LGD_REQUIRE(count, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) { *count = 0; return 0; }
    LGD_HIP(e);
    *count = n;
    return 0;
}

int lgdhip_set_device(int device) {
    LGD_HIP(hipSetDevice(device));
    return 0;
}

int lgdhip_get_device(int *device) {
    LGD_REQUIRE(device, "NULL pointer");
    LGD_HIP(hipGetDevice(device));
    return 0;
}

int lgdhip_device_name(int device, char *buf, int len) {
    LGD_REQUIRE(buf && len > 0, "bad output buffer");
    hipDeviceProp_t prop;
    LGD_HIP(hipGetDeviceProperties(&prop, device));
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return 0;
}

int lgdhip_device_synchronize(void) {
    LGD_HIP(hipDeviceSynchronize());
    return 0;
}

int lgdhip_malloc(void **ptr, size_t bytes) {
    LGD_REQUIRE(ptr, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return 0;
    LGD_HIP(hipMalloc(ptr, bytes));
    return 0;
}

int lgdhip_free(void *ptr) {
    if (!ptr) return 0;
    LGD_HIP(hipFree(ptr));
    return 0;
}

int lgdhip_host_malloc(void **ptr, size_t bytes) {
    LGD_REQUIRE(ptr, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return 0;
    LGD_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return 0;
}

int lgdhip_host_free(void *ptr) {
    if (!ptr) return 0;
    LGD_HIP(hipHostFree(ptr));
    return 0;
}

int lgdhip_memcpy_h2d(void *dst, const void *src, size_t bytes, lgdhip_stream_t stream) {
    if (bytes == 0) return 0;
    LGD_REQUIRE(dst && src, "NULL pointer");
    LGD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    // the source may be pageable caller memory: finish before returning
    LGD_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int lgdhip_memcpy_d2h(void *dst, const void *src, size_t bytes, lgdhip_stream_t stream) {
    if (bytes == 0) return 0;
    LGD_REQUIRE(dst && src, "NULL pointer");
    LGD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    LGD_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int lgdhip_memcpy2d_d2h(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                        lgdhip_stream_t stream) {
    if (width == 0 || height == 0) return 0;
    LGD_REQUIRE(dst && src && dpitch >= width && spitch >= width, "NULL pointer or pitch below the row width");
    LGD_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost, as_stream(stream)));
    LGD_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int lgdhip_memcpy_d2d(void *dst, const void *src, size_t bytes, lgdhip_stream_t stream) {
    if (bytes == 0) return 0;
    LGD_REQUIRE(dst && src, "NULL pointer");
    LGD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return 0;
}

int lgdhip_fill_u8(void *dst, uint8_t value, size_t bytes, lgdhip_stream_t stream) {
    if (bytes == 0) return 0;
    LGD_REQUIRE(dst, "NULL pointer");
    if (((uintptr_t)dst & 3) == 0 && (bytes & 3) == 0) {
        const uint32_t v = 0x01010101u * value;
        lgd_fill_u32_kernel<<<fill_blocks(bytes / 4), 256, 0, as_stream(stream)>>>((uint32_t *)dst, v, bytes / 4);
    } else {
        lgd_fill_u8_kernel<<<fill_blocks(bytes), 256, 0, as_stream(stream)>>>((uint8_t *)dst, value, bytes);
    }
    LGD_LAUNCHED("lgdhip_fill_u8");
    return 0;
}

int lgdhip_copy_kernel(void *dst, const void *src, size_t bytes, lgdhip_stream_t stream) {
    LGD_REQUIRE(dst && src, "NULL pointer");
    LGD_REQUIRE(bytes % 16 == 0, "bytes must be a multiple of 16");
    LGD_REQUIRE(((uintptr_t)dst & 15) == 0 && ((uintptr_t)src & 15) == 0, "pointers must be 16-byte aligned");
    const size_t n = bytes / 16;
    if (n == 0) return 0;
    const size_t blocks = (n + 255) / 256;
    LGD_REQUIRE(blocks <= 0xFFFFFFFFull, "copy too large for one launch");
    lgd_copy16_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>((uint4 *)dst, (const uint4 *)src, n);
    LGD_LAUNCHED("lgdhip_copy_kernel");
    return 0;
}

int lgdhip_stream_create(lgdhip_stream_t *stream) {
    LGD_REQUIRE(stream, "stream is NULL");
    hipStream_t s;
    LGD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (lgdhip_stream_t)s;
    return 0;
}

int lgdhip_stream_destroy(lgdhip_stream_t stream) {
    if (!stream) return 0;
    LGD_HIP(hipStreamDestroy(as_stream(stream)));
    return 0;
}

int lgdhip_stream_synchronize(lgdhip_stream_t stream) {
    LGD_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int lgdhip_event_create(lgdhip_event_t *event) {
    LGD_REQUIRE(event, "event is NULL");
    hipEvent_t e;
    LGD_HIP(hipEventCreate(&e));
    *event = (lgdhip_event_t)e;
    return 0;
}

int lgdhip_event_destroy(lgdhip_event_t event) {
    if (!event) return 0;
    LGD_HIP(hipEventDestroy((hipEvent_t)event));
    return 0;
}

int lgdhip_event_record(lgdhip_event_t event, lgdhip_stream_t stream) {
    LGD_REQUIRE(event, "event is NULL");
    LGD_HIP(hipEventRecord((hipEvent_t)event, as_stream(stream)));
    return 0;
}

int lgdhip_stream_wait_event(lgdhip_stream_t stream, lgdhip_event_t event) {
    LGD_REQUIRE(event, "event is NULL");
    LGD_HIP(hipStreamWaitEvent(as_stream(stream), (hipEvent_t)event, 0));
    return 0;
}

int lgdhip_event_synchronize(lgdhip_event_t event) {
    LGD_REQUIRE(event, "event is NULL");
    LGD_HIP(hipEventSynchronize((hipEvent_t)event));
    return 0;
}

int lgdhip_event_elapsed_ms(float *ms, lgdhip_event_t start, lgdhip_event_t stop) {
    LGD_REQUIRE(ms && start && stop, "NULL argument");
    LGD_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return 0;
}

int lgdhip_graph_begin(lgdhip_stream_t stream) {
    LGD_REQUIRE(stream, "graph capture needs a non-null stream");
    LGD_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
    return 0;
}

int lgdhip_graph_end(lgdhip_stream_t stream, lgdhip_graph_t *graph) {
    LGD_REQUIRE(stream && graph, "NULL argument");
    hipGraph_t g = nullptr;
    LGD_HIP(hipStreamEndCapture(as_stream(stream), &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph = (lgdhip_graph_t)ex;
    return 0;
}

int lgdhip_graph_launch(lgdhip_graph_t graph, lgdhip_stream_t stream) {
    LGD_REQUIRE(graph, "graph is NULL");
    LGD_HIP(hipGraphLaunch((hipGraphExec_t)graph, as_stream(stream)));
    return 0;
}

int lgdhip_graph_destroy(lgdhip_graph_t graph) {
    if (!graph) return 0;
    LGD_HIP(hipGraphExecDestroy((hipGraphExec_t)graph));
    return 0;
}

// ---------------------------------------------------------------------------
// grids
// ---------------------------------------------------------------------------
int lgdhip_grid_make(lgdhip_grid *g, int dim, int64_t nx, int64_t ny, int64_t nz, int halo,
                     int elem_bytes, int members) {
    LGD_REQUIRE(g, "grid descriptor is NULL");
    LGD_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    LGD_REQUIRE(halo >= 1 && halo <= 64, "halo must be in [1, 64]");
    LGD_REQUIRE(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8,
                "elem_bytes must be 1, 2, 4 or 8");
    LGD_REQUIRE(members >= 1 && members <= 64, "members must be in [1, 64]");
    if (dim == 2) nz = 1;
    LGD_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "extents must be positive");
    std::memset(g, 0, sizeof(*g));
    g->dim = dim;
    g->halo = halo;
    g->elem_bytes = elem_bytes;
    g->members = members;
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    // the first logical cell of every row starts on a 128-byte boundary and
    // the row pitch is a multiple of 128 bytes (one gfx950 cache line), so
    // wave-wide row accesses cover whole lines; at least 64 bytes of padding
    // follow the last cell so vector kernels may read whole groups past nx
    const int64_t per64 = 64 / elem_bytes;
    const int64_t per128 = 128 / elem_bytes;
    const int64_t right = halo > per64 ? halo : per64;
    g->ox = ((halo + per128 - 1) / per128) * per128;
    g->sx = ((g->ox + nx + right + per128 - 1) / per128) * per128;
    g->oy = halo;
    g->sy = ny + 2 * halo;
    if (dim == 3) {
        g->oz = halo;
        g->sz = nz + 2 * halo;
    } else {
        g->oz = 0;
        g->sz = 1;
    }
    return check_grid(g, dim, elem_bytes, members);
}

int lgdhip_grid_view(const lgdhip_grid *g, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny,
                     int64_t nz, lgdhip_grid *view) {
    int rc = check_grid(g, 0, 0, 0, true);
    if (rc) return rc;
    LGD_REQUIRE(view, "view descriptor is NULL");
    if (g->dim == 2) {
        LGD_REQUIRE(z0 == 0 && nz == 1, "a view of a 2D grid has z0 == 0 and nz == 1");
    }
    LGD_REQUIRE(x0 >= 0 && y0 >= 0 && z0 >= 0 && nx >= 1 && ny >= 1 && nz >= 1, "empty or negative view box");
    LGD_REQUIRE(x0 + nx <= g->nx && y0 + ny <= g->ny && z0 + nz <= g->nz, "view box outside the grid");
    lgdhip_grid v = *g;
    v.ox += x0;
    v.oy += y0;
    v.oz += z0;
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    v.flags |= LGDHIP_GRID_VIEW;
    *view = v;
    return check_grid(view, 0, 0, 0, true);
}

int lgdhip_grid_set_layout(lgdhip_grid *g, int layout) {
    LGD_REQUIRE(g, "grid descriptor is NULL");
    LGD_REQUIRE(layout == LGDHIP_LAYOUT_MEMBER_PLANES || layout == LGDHIP_LAYOUT_Z_SLABS, "unknown grid layout");
    g->layout = layout;
    return check_grid(g, 0, 0, 0);
}

int64_t lgdhip_grid_plane_elems(const lgdhip_grid *g) { return g ? plane_elems(*g) : 0; }

int64_t lgdhip_grid_bytes(const lgdhip_grid *g) {
    return g ? plane_elems(*g) * g->members * g->elem_bytes : 0;
}

int lgdhip_grid_fill_halo(void *grid, const lgdhip_grid *g, const void *edge, int wrap_mask,
                          lgdhip_stream_t stream) {
    int rc = check_grid(g, 0, 0, 0);
    if (rc) return rc;
    LGD_REQUIRE(grid && edge, "NULL pointer");
    LGD_REQUIRE(wrap_mask >= 0 && wrap_mask <= 7, "wrap_mask must be a combination of bits 1, 2, 4");
    const int64_t n = member_stride(*g);
    const int h = g->halo;
    const int64_t shell = (g->nx + 2 * h) * (g->ny + 2 * h) * (g->dim == 3 ? g->nz + 2 * h : 1);
    const unsigned blocks = fill_blocks((size_t)shell);
    hipStream_t s = as_stream(stream);
    const int64_t zs = z_stride(*g);
    for (int m = 0; m < g->members; ++m) {
        const uint8_t *ev = (const uint8_t *)edge + (size_t)m * g->elem_bytes;
        char *plane = (char *)grid + (size_t)m * n * g->elem_bytes;
#define LGD_HALO_CASE(BYTES, T)                                                                                   \
    case BYTES: {                                                                                                 \
        T v;                                                                                                      \
        std::memcpy(&v, ev, BYTES);                                                                               \
        lgd_fill_halo_kernel<T><<<blocks, 256, 0, s>>>((T *)plane, v, g->sx, zs, g->ox, g->oy, g->oz, g->nx,     \
                                                       g->ny, g->nz, h, g->dim, wrap_mask);                       \
        break;                                                                                                    \
    }
        switch (g->elem_bytes) {
            LGD_HALO_CASE(1, uint8_t)
            LGD_HALO_CASE(2, uint16_t)
            LGD_HALO_CASE(4, uint32_t)
            LGD_HALO_CASE(8, uint64_t)
        default: return set_error(LGDHIP_ERR_UNSUPPORTED, "element size must be 1, 2, 4 or 8 bytes");
        }
#undef LGD_HALO_CASE
        LGD_LAUNCHED("lgdhip_grid_fill_halo");
    }
    return 0;
}

int lgdhip_grid_fill_edge(void *grid, const lgdhip_grid *g, const void *edge, lgdhip_stream_t stream) {
    int rc = check_grid(g, 0, 0, 0);
    if (rc) return rc;
    LGD_REQUIRE(grid && edge, "NULL pointer");
    const unsigned blocks = fill_blocks((size_t)plane_elems(*g));
    const int64_t ms = member_stride(*g), zs = z_stride(*g);
    for (int m = 0; m < g->members; ++m) {
        const uint8_t *ev = (const uint8_t *)edge + (size_t)m * g->elem_bytes;
        char *plane = (char *)grid + (size_t)m * ms * g->elem_bytes;
        switch (g->elem_bytes) {
        case 1:
            lgd_fill_edge_kernel<uint8_t><<<blocks, 256, 0, as_stream(stream)>>>(
                (uint8_t *)plane, *ev, g->sx, g->sy, g->sz, zs, g->ox, g->oy, g->oz, g->nx, g->ny, g->nz);
            break;
        case 2: {
            uint16_t v;
            std::memcpy(&v, ev, 2);
            lgd_fill_edge_kernel<uint16_t><<<blocks, 256, 0, as_stream(stream)>>>(
                (uint16_t *)plane, v, g->sx, g->sy, g->sz, zs, g->ox, g->oy, g->oz, g->nx, g->ny, g->nz);
            break;
        }
        case 4: {
            uint32_t v;
            std::memcpy(&v, ev, 4);
            lgd_fill_edge_kernel<uint32_t><<<blocks, 256, 0, as_stream(stream)>>>(
                (uint32_t *)plane, v, g->sx, g->sy, g->sz, zs, g->ox, g->oy, g->oz, g->nx, g->ny, g->nz);
            break;
        }
        default: {
            uint64_t v;
            std::memcpy(&v, ev, 8);
            lgd_fill_edge_kernel<uint64_t><<<blocks, 256, 0, as_stream(stream)>>>(
                (uint64_t *)plane, v, g->sx, g->sy, g->sz, zs, g->ox, g->oy, g->oz, g->nx, g->ny, g->nz);
            break;
        }
        }
        LGD_LAUNCHED("lgdhip_grid_fill_edge");
    }
    return 0;
}

static int grid_copy(void *dev, const lgdhip_grid *g, void *host, bool to_device, lgdhip_stream_t stream) {
    int rc = check_grid(g, 0, 0, 0, true);
    if (rc) return rc;
    LGD_REQUIRE(dev && host, "NULL pointer");
    const size_t eb = (size_t)g->elem_bytes;
    const int64_t n = member_stride(*g);
    // rows per z slice of the device side: sy, or members * sy for z slabs
    const size_t slice_rows = (size_t)(z_stride(*g) / g->sx);
    const size_t host_plane = (size_t)(g->nx * g->ny * g->nz) * eb;
    for (int m = 0; m < g->members; ++m) {
        char *dplane = (char *)dev + (size_t)m * n * eb;
        char *hplane = (char *)host + (size_t)m * host_plane;
        hipMemcpy3DParms p;
        std::memset(&p, 0, sizeof(p));
        hipPitchedPtr dptr = make_hipPitchedPtr(dplane, (size_t)g->sx * eb, (size_t)g->sx * eb, slice_rows);
        hipPitchedPtr hptr = make_hipPitchedPtr(hplane, (size_t)g->nx * eb, (size_t)g->nx * eb, (size_t)g->ny);
        hipPos dpos = make_hipPos((size_t)g->ox * eb, (size_t)g->oy, (size_t)g->oz);
        hipPos hpos = make_hipPos(0, 0, 0);
        p.extent = make_hipExtent((size_t)g->nx * eb, (size_t)g->ny, (size_t)g->nz);
        if (to_device) {
            p.srcPtr = hptr;
            p.srcPos = hpos;
            p.dstPtr = dptr;
            p.dstPos = dpos;
            p.kind = hipMemcpyHostToDevice;
        } else {
            p.srcPtr = dptr;
            p.srcPos = dpos;
            p.dstPtr = hptr;
            p.dstPos = hpos;
            p.kind = hipMemcpyDeviceToHost;
        }
        LGD_HIP(hipMemcpy3DAsync(&p, as_stream(stream)));
    }
    // the host buffer belongs to the caller and may be released right after
    // this call returns, so both directions complete before returning
    LGD_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int lgdhip_grid_load(void *grid, const lgdhip_grid *g, const void *host, lgdhip_stream_t stream) {
    return grid_copy(grid, g, const_cast<void *>(host), true, stream);
}

int lgdhip_grid_store(const void *grid, const lgdhip_grid *g, void *host, lgdhip_stream_t stream) {
    return grid_copy(const_cast<void *>(grid), g, host, false, stream);
}

} // extern "C"
