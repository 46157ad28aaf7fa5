// Cell grids: device array-of-structs storage of whole user cells, the
// storage of the generic kernel family (user cells whose update() is
// compiled by hipcc).  Padding, halo refresh for Cube / Torus / mixed
// topologies, and host <-> device copies of the logical box.
//
// Replaces the padded AoS device buffers of CUDASimulator
// (src/libgeodecomp/parallelization/cudasimulator.h:314-429, :502-534):
// there the padding holds the edge cell once and wrapped axes are handled
// by per-thread index offsets in the kernel (:52-64); here every axis is
// padded by the stencil radius and wrapped axes are refreshed by one halo
// kernel before each sweep, so the sweep kernel reads neighbours without
// branches.
#include "lgdhip_internal.h"

#include <cstring>

using namespace lgdhip;

namespace {

struct ShellShape {
    int64_t nx, ny, nz, h, hy, hz, px, py;
    int64_t a, b, c;   // cells in the z-halo planes, the y-halo rows, the x-halo columns
};

ShellShape shell_of(const lgdhip_cell_grid &g) {
    ShellShape s;
    s.nx = g.nx;
    s.ny = g.ny;
    s.nz = g.nz;
    s.h = g.halo;
    s.hy = g.dim >= 2 ? g.halo : 0;
    s.hz = g.dim == 3 ? g.halo : 0;
    s.px = g.px;
    s.py = g.py;
    s.a = 2 * s.hz * g.px * g.py;
    s.b = g.nz * 2 * s.hy * g.px;
    s.c = g.nz * g.ny * 2 * s.h;
    return s;
}

// Halo cell `i` of the shell (z-halo planes first, then the y-halo rows of
// the interior planes, then the x-halo cells of the interior rows) -> its
// logical coordinate (may be negative or >= the extent).
__device__ inline void shell_coord(const ShellShape &s, int64_t i, int64_t &x, int64_t &y, int64_t &z) {
    if (i < s.a) {
        const int64_t plane = i / (s.px * s.py), r = i % (s.px * s.py);
        z = plane < s.hz ? plane - s.hz : s.nz + (plane - s.hz);
        y = r / s.px - s.hy;
        x = r % s.px - s.h;
        return;
    }
    i -= s.a;
    if (i < s.b) {
        const int64_t rows = 2 * s.hy;
        z = i / (rows * s.px);
        const int64_t r = i % (rows * s.px), row = r / s.px;
        y = row < s.hy ? row - s.hy : s.ny + (row - s.hy);
        x = r % s.px - s.h;
        return;
    }
    i -= s.b;
    const int64_t cols = 2 * s.h;
    z = i / (s.ny * cols);
    const int64_t r = i % (s.ny * cols), col = r % cols;
    y = r / cols;
    x = col < s.h ? col - s.h : s.nx + (col - s.h);
}

__device__ inline int64_t wrap(int64_t v, int64_t n) { return ((v % n) + n) % n; }

// One thread per (halo cell, word of the cell).  A halo cell copies the
// interior cell at its coordinate wrapped along the axes in `wrap_mask`, or
// the edge cell when the coordinate is still outside an unwrapped axis (the
// neighbourhood rule of the reference's topologies: out of bounds first,
// then normalise, geometry/topologies.h:185-234).
template<typename W>
__global__ void __launch_bounds__(256) cell_fill_halo_kernel(W *__restrict__ cells, const W *__restrict__ edge,
                                                             ShellShape s, int words, int wrap_mask) {
    const int64_t total = (s.a + s.b + s.c) * words;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / words;
        const int w = (int)(t % words);
        int64_t x, y, z;
        shell_coord(s, i, x, y, z);
        int64_t wx = x, wy = y, wz = z;
        if (wrap_mask & 1) wx = wrap(x, s.nx);
        if (wrap_mask & 2) wy = wrap(y, s.ny);
        if (wrap_mask & 4) wz = wrap(z, s.nz);
        const bool inside = wx >= 0 && wx < s.nx && wy >= 0 && wy < s.ny && wz >= 0 && wz < s.nz;
        const int64_t dst = ((z + s.hz) * s.py + (y + s.hy)) * s.px + (x + s.h);
        // outside: point the (unused) source at the destination, so the
        // index is in the buffer whatever the compiler does with the select
        const int64_t src = inside ? ((wz + s.hz) * s.py + (wy + s.hy)) * s.px + (wx + s.h) : dst;
        cells[dst * words + w] = inside ? cells[src * words + w] : edge[w];
    }
}

int check_cell_grid(const lgdhip_cell_grid *g) {
    LGD_REQUIRE(g, "cell grid descriptor is NULL");
    LGD_REQUIRE(g->dim >= 1 && g->dim <= 3, "dim must be 1, 2 or 3");
    LGD_REQUIRE(g->halo >= 1 && g->halo <= 16, "halo must be in [1, 16]");
    LGD_REQUIRE(g->cell_bytes >= 1 && g->cell_bytes <= 65536, "cell_bytes must be in [1, 65536]");
    LGD_REQUIRE(g->nx >= 1 && g->ny >= 1 && g->nz >= 1, "extents must be positive");
    LGD_REQUIRE(g->dim >= 2 || g->ny == 1, "a 1D grid has ny == 1");
    LGD_REQUIRE(g->dim == 3 || g->nz == 1, "a 1D or 2D grid has nz == 1");
    LGD_REQUIRE(g->px == g->nx + 2 * g->halo, "px must be nx + 2 halo");
    LGD_REQUIRE(g->py == (g->dim >= 2 ? g->ny + 2 * g->halo : 1), "py must be ny + 2 halo (1 in 1D)");
    LGD_REQUIRE(g->pz == (g->dim == 3 ? g->nz + 2 * g->halo : 1), "pz must be nz + 2 halo (1 below 3D)");
    return 0;
}

int cell_copy(void *dev, const lgdhip_cell_grid *g, void *host, bool to_device, lgdhip_stream_t stream) {
    int rc = check_cell_grid(g);
    if (rc) return rc;
    LGD_REQUIRE(dev && host, "NULL pointer");
    const size_t cb = (size_t)g->cell_bytes;
    const int64_t hy = g->dim >= 2 ? g->halo : 0, hz = g->dim == 3 ? g->halo : 0;
    hipMemcpy3DParms p;
    std::memset(&p, 0, sizeof(p));
    hipPitchedPtr dptr = make_hipPitchedPtr(dev, (size_t)g->px * cb, (size_t)g->px * cb, (size_t)g->py);
    hipPitchedPtr hptr = make_hipPitchedPtr(host, (size_t)g->nx * cb, (size_t)g->nx * cb, (size_t)g->ny);
    hipPos dpos = make_hipPos((size_t)g->halo * cb, (size_t)hy, (size_t)hz);
    hipPos hpos = make_hipPos(0, 0, 0);
    p.extent = make_hipExtent((size_t)g->nx * cb, (size_t)g->ny, (size_t)g->nz);
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
    // the host buffer belongs to the caller: complete before returning
    LGD_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

} // namespace

extern "C" {

int lgdhip_cell_grid_make(lgdhip_cell_grid *g, int dim, int64_t nx, int64_t ny, int64_t nz, int halo,
                          int64_t cell_bytes) {
    LGD_REQUIRE(g, "cell grid descriptor is NULL");
    LGD_REQUIRE(dim >= 1 && dim <= 3, "dim must be 1, 2 or 3");
    if (dim < 3) nz = 1;
    if (dim < 2) ny = 1;
    std::memset(g, 0, sizeof(*g));
    g->dim = dim;
    g->halo = halo;
    g->cell_bytes = cell_bytes;
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    g->px = nx + 2 * (int64_t)halo;
    g->py = dim >= 2 ? ny + 2 * (int64_t)halo : 1;
    g->pz = dim == 3 ? nz + 2 * (int64_t)halo : 1;
    return check_cell_grid(g);
}

int64_t lgdhip_cell_grid_bytes(const lgdhip_cell_grid *g) {
    if (check_cell_grid(g)) return 0;
    return g->px * g->py * g->pz * g->cell_bytes;
}

int lgdhip_cell_grid_fill_halo(void *cells, const lgdhip_cell_grid *g, const void *edge, int wrap_mask,
                               lgdhip_stream_t stream) {
    int rc = check_cell_grid(g);
    if (rc) return rc;
    LGD_REQUIRE(cells && edge, "NULL pointer");
    LGD_REQUIRE(wrap_mask >= 0 && wrap_mask < (1 << g->dim), "wrap_mask names an axis the grid does not have");
    const ShellShape s = shell_of(*g);
    const bool word4 = g->cell_bytes % 4 == 0 && (uintptr_t)cells % 4 == 0 && (uintptr_t)edge % 4 == 0;
    const int words = (int)(word4 ? g->cell_bytes / 4 : g->cell_bytes);
    const int64_t total = (s.a + s.b + s.c) * words;
    if (total == 0) return 0;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipStream_t st = as_stream(stream);
    if (word4) {
        cell_fill_halo_kernel<uint32_t><<<(unsigned)blocks, 256, 0, st>>>((uint32_t *)cells, (const uint32_t *)edge,
                                                                          s, words, wrap_mask);
    } else {
        cell_fill_halo_kernel<uint8_t><<<(unsigned)blocks, 256, 0, st>>>((uint8_t *)cells, (const uint8_t *)edge,
                                                                         s, words, wrap_mask);
    }
    LGD_LAUNCHED("lgdhip_cell_grid_fill_halo");
    return 0;
}

int lgdhip_cell_grid_load(void *cells, const lgdhip_cell_grid *g, const void *host, lgdhip_stream_t stream) {
    return cell_copy(cells, g, const_cast<void *>(host), true, stream);
}

int lgdhip_cell_grid_store(const void *cells, const lgdhip_cell_grid *g, void *host, lgdhip_stream_t stream) {
    return cell_copy(const_cast<void *>(cells), g, host, false, stream);
}

} // extern "C"
