// HipKernels::Generic<CELL>: the kernel family of every cell that has no
// hand-written family -- the user's own update() compiled for gfx950.
//
// Include this header from a translation unit compiled by hipcc
// (--offload-arch=gfx950); HipSimulator<CELL> then runs such cells without a
// registration line, like the reference's CUDASimulator runs any cell whose
// update() is __host__ __device__ (cudasimulator.h:25-290).  The cell is
// stored as whole structs (array of structs) in a cell grid of the lgdhip
// C-ABI (lgdhip_cell_grid), padded on every axis by the stencil radius.
// Per nano step:
//   1. wrapped axes (Torus and mixed topologies) refresh their halo from the
//      opposite side (lgdhip_cell_grid_fill_halo); unwrapped halos hold the
//      edge cell, written at upload (the reference pads once and wraps with
//      per-thread index offsets, cudasimulator.h:52-64);
//   2. one thread per cell copies its target cell into registers, calls
//      update(hood, nanoStep) -- or updateHIP() / updateCUDA() under
//      APITraits::HasSeparateHIPUpdate / HasSeparateCUDAUpdate, as
//      storage/cudaupdatefunctor.h:13-42 selects -- and stores it back;
//      hood[FixedCoord<X,Y,Z>()] and hood[Coord<DIM>] are plain loads at a
//      fixed offset from the cell, no bounds logic;
//   3. the two buffers swap.
// The target cell starts as the state of two nano steps ago, exactly as in
// the reference's double-buffered simulators (cells that do not overwrite
// all their members see the same values on the GPU as on the CPU).
//
// Requirements on the cell: update() and what it calls are
// __host__ __device__ (libgeodecomp/misc/hostdevice.h makes the attributes
// empty for host compilers); it reads only neighbours within the stencil
// radius (APITraits::HasStencil, default Moore<DIM, 1>); the cell is
// trivially copyable (it is moved between host and device as bytes).
#ifndef LIBGEODECOMP_HIP_GENERICFAMILY_H
#define LIBGEODECOMP_HIP_GENERICFAMILY_H

#ifndef __HIP__
#error "libgeodecomp/hip/genericfamily.h compiles user cells for the GPU: build this translation unit with hipcc"
#endif

#include <hip/hip_runtime.h>

#include <libgeodecomp/geometry/coord.h>
#include <libgeodecomp/geometry/fixedcoord.h>
#include <libgeodecomp/hip/kernelfamilies.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace LibGeoDecomp {

namespace HipKernels {

// Neighbourhood of one cell in a padded cell grid: element offsets 1, px,
// px * py along x, y, z.
template<typename CELL, int DIM>
class GenericNeighborhood
{
public:
    __device__ GenericNeighborhood(const CELL *center, long px, long pxy) : center(center)
    {
        stride[0] = 1;
        stride[1] = px;
        stride[2] = pxy;
    }

    template<int X, int Y, int Z>
    __device__ const CELL& operator[](FixedCoord<X, Y, Z>) const
    {
        return center[X + (DIM > 1 ? Y * stride[1] : 0) + (DIM > 2 ? Z * stride[2] : 0)];
    }

    __device__ const CELL& operator[](const Coord<DIM>& rel) const
    {
        long off = 0;
        for (int d = 0; d < DIM; ++d) {
            off += rel[d] * stride[d];
        }
        return center[off];
    }

private:
    const CELL *center;
    long stride[3];
};

namespace GenericHelpers {

// update() / updateHIP() / updateCUDA(), as the cell's API selects
template<typename CELL, typename SEPARATE_HIP, typename SEPARATE_CUDA>
class Update
{
public:
    template<typename HOOD>
    __device__ static void apply(CELL& cell, const HOOD& hood, unsigned nanoStep) { cell.update(hood, nanoStep); }
};

template<typename CELL, typename SEPARATE_CUDA>
class Update<CELL, APITraits::TrueType, SEPARATE_CUDA>
{
public:
    template<typename HOOD>
    __device__ static void apply(CELL& cell, const HOOD& hood, unsigned nanoStep)
    {
        cell.updateHIP(hood, nanoStep);
    }
};

template<typename CELL>
class Update<CELL, APITraits::FalseType, APITraits::TrueType>
{
public:
    template<typename HOOD>
    __device__ static void apply(CELL& cell, const HOOD& hood, unsigned nanoStep)
    {
        cell.updateCUDA(hood, nanoStep);
    }
};

// Neighbourhood of an array-of-structs line update (a cell with
// HasFixedCoordsOnlyUpdate and HasUpdateLineX but no SoA):
// hood[FixedCoord<X,Y,Z>()] is the cell at (*x + X, y + Y, z + Z) of the
// row the update walks, x shared with the cell's loop, as the reference's
// LinePointerNeighborhood reads its line pointers at the current x
// (storage/linepointerneighborhood.h).
template<typename CELL, int DIM>
class GenericLineNeighborhood
{
public:
    __device__ GenericLineNeighborhood(const CELL *row, const long *x, long px, long pxy) : row(row), x(x)
    {
        stride[0] = 1;
        stride[1] = px;
        stride[2] = pxy;
    }

    template<int X, int Y, int Z>
    __device__ const CELL& operator[](FixedCoord<X, Y, Z>) const
    {
        return row[*x + X + (DIM > 1 ? Y * stride[1] : 0) + (DIM > 2 ? Z * stride[2] : 0)];
    }

private:
    const CELL *row;
    const long *x;
    long stride[3];
};

// update of the cell at index i (row start `row`, x coordinate x): the
// cell-wise update() path, or one element of a line update
template<typename CELL, int DIM, bool LINE = APITraits::SelectLinePointerUpdate<CELL>::VALUE>
class UpdateAt
{
public:
    __device__ static void apply(const CELL *src, CELL *dst, long row, int x, long px, long pxy, unsigned nanoStep)
    {
        typedef Update<CELL, typename APITraits::SelectSeparateHIPUpdate<CELL>::Value,
                       typename APITraits::SelectSeparateCUDAUpdate<CELL>::Value> U;
        const long i = row + x;
        GenericNeighborhood<CELL, DIM> hood(src + i, px, pxy);
        CELL cell = dst[i];
        U::apply(cell, hood, nanoStep);
        dst[i] = cell;
    }
};

template<typename CELL, int DIM>
class UpdateAt<CELL, DIM, true>
{
public:
    __device__ static void apply(const CELL *src, CELL *dst, long row, int x, long px, long pxy, unsigned nanoStep)
    {
        long cursor = x;
        GenericLineNeighborhood<CELL, DIM> hood(src + row, &cursor, px, pxy);
        CELL::updateLineX(dst + row, &cursor, (long)x + 1, hood, nanoStep);
    }
};

// One thread per cell.  Work-group (bx, by): bx consecutive x cells of by
// consecutive rows (a row = one (y, z) pair); blockIdx.x enumerates
// (row group, x block) with the x block fastest.
template<typename CELL, int DIM>
__global__ void __launch_bounds__(1024) sweep(const CELL *__restrict__ src, CELL *__restrict__ dst, int nx,
                                              long rows, int ny, int h, int hy, int hz, long px, long pxy,
                                              unsigned xblocks, unsigned nanoStep)
{
    const unsigned xb = blockIdx.x % xblocks;
    const long row = (long)(blockIdx.x / xblocks) * blockDim.y + threadIdx.y;
    const int x = (int)(xb * blockDim.x + threadIdx.x);
    if (x >= nx || row >= rows) {
        return;
    }
    const long y = row % ny;
    const long z = row / ny;
    UpdateAt<CELL, DIM>::apply(src, dst, (z + hz) * pxy + (y + hy) * px + h, x, px, pxy, nanoStep);
}

inline void checkHip(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        throw std::runtime_error(std::string(what) + " failed: " + hipGetErrorString(e));
    }
}

}

template<typename CELL>
class Generic
{
public:
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    typedef typename APITraits::SelectStencil<CELL>::Value Stencil;
    static const int DIM = Topology::DIM;
    static const int RADIUS = Stencil::RADIUS;
    static const unsigned NANO_STEPS = APITraits::SelectNanoSteps<CELL>::VALUE;

    static_assert(std::is_trivially_copyable<CELL>::value,
                  "the generic HIP kernel family moves cells between host and device as bytes: the cell must be "
                  "trivially copyable");
    static_assert(RADIUS >= 1 && RADIUS <= 16, "stencil radius must be in [1, 16]");

    // tuning.tile.bx / by: work-group shape (x cells, rows); zero = 64 x 4
    // (2D, 3D) or 256 x 1 (1D)
    Generic(const Coord<DIM>& dim, const Tuning& tuning) :
        wrap(wrapMask<CELL>()),
        uploaded(false),
        bx(DIM == 1 ? 256 : 64),
        by(DIM == 1 ? 1 : 4)
    {
        if (tuning.useTile && tuning.tile.bx > 0) {
            bx = tuning.tile.bx;
        }
        if (tuning.useTile && tuning.tile.by > 0) {
            by = tuning.tile.by;
        }
        if (bx * by > 1024 || bx % 64 != 0) {
            throw std::invalid_argument("generic HIP kernel: work-group width must be a multiple of 64 and "
                                        "width x rows at most 1024");
        }
        const long ny = DIM >= 2 ? dim[DIM >= 2 ? 1 : 0] : 1;
        const long nz = DIM == 3 ? dim[DIM == 3 ? 2 : 0] : 1;
        check(lgdhip_cell_grid_make(&g, DIM, dim[0], ny, nz, RADIUS, (int64_t)sizeof(CELL)), "lgdhip_cell_grid_make");
        const std::size_t bytes = (std::size_t)lgdhip_cell_grid_bytes(&g);
        a.reset(new DeviceMemory(bytes));
        b.reset(new DeviceMemory(bytes));
        edge.reset(new DeviceMemory(sizeof(CELL)));
    }

    void upload(const Grid<CELL, Topology>& host)
    {
        requireBox(host);
        const CELL e = host.getEdge();
        check(lgdhip_memcpy_h2d(edge->get(), &e, sizeof(CELL), stream.get()), "lgdhip_memcpy_h2d");
        check(lgdhip_cell_grid_load(a->get(), &g, host.data(), stream.get()), "lgdhip_cell_grid_load");
        if (!uploaded) {
            // both buffers start from the initial grid, as the reference's
            // simulators initialise both grids
            check(lgdhip_cell_grid_load(b->get(), &g, host.data(), stream.get()), "lgdhip_cell_grid_load");
            uploaded = true;
        }
        check(lgdhip_cell_grid_fill_halo(a->get(), &g, edge->get(), wrap, stream.get()), "lgdhip_cell_grid_fill_halo");
        check(lgdhip_cell_grid_fill_halo(b->get(), &g, edge->get(), wrap, stream.get()), "lgdhip_cell_grid_fill_halo");
    }

    void uploadRegion(const Grid<CELL, Topology>& host, const Region<DIM>& region)
    {
        requireBox(host);
        const std::vector<Streak<DIM> > streaks = relativeStreaks(region, host.boundingBox());
        const long hy = DIM >= 2 ? RADIUS : 0, hz = DIM == 3 ? RADIUS : 0;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<DIM>& st = streaks[k];
            const long y = coordOf(st.origin, 1), z = coordOf(st.origin, 2);
            const long i = ((z + hz) * g.py + (y + hy)) * g.px + (st.origin.x() + RADIUS);
            const CELL *src = host.data() + st.origin.toIndex(host.boundingBox().dimensions);
            check(lgdhip_memcpy_h2d(static_cast<CELL *>(a->get()) + i, src,
                                    sizeof(CELL) * (std::size_t)(st.endX - st.origin.x()), stream.get()),
                  "lgdhip_memcpy_h2d");
        }
    }

    void download(Grid<CELL, Topology> *host)
    {
        requireBox(*host);
        check(lgdhip_cell_grid_store(a->get(), &g, host->data(), stream.get()), "lgdhip_cell_grid_store");
    }

    // one contiguous copy per streak straight into the host cells
    void downloadRegion(Grid<CELL, Topology> *host, const Region<DIM>& region)
    {
        requireBox(*host);
        const std::vector<Streak<DIM> > streaks = relativeStreaks(region, host->boundingBox());
        const long hy = DIM >= 2 ? RADIUS : 0, hz = DIM == 3 ? RADIUS : 0;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<DIM>& st = streaks[k];
            const long y = coordOf(st.origin, 1), z = coordOf(st.origin, 2);
            const long i = ((z + hz) * g.py + (y + hy)) * g.px + (st.origin.x() + RADIUS);
            CELL *dst = host->data() + st.origin.toIndex(host->boundingBox().dimensions);
            check(lgdhip_memcpy_d2h(dst, static_cast<const CELL *>(a->get()) + i,
                                    sizeof(CELL) * (std::size_t)(st.endX - st.origin.x()), stream.get()),
                  "lgdhip_memcpy_d2h");
        }
    }

    // nanoSteps is a multiple of NANO_STEPS (whole steps).  A period of
    // lcm(2, NANO_STEPS) nano steps returns the buffers to their places and
    // is replayed as one captured graph; the rest is launched directly.
    void run(unsigned nanoSteps)
    {
        const unsigned period = NANO_STEPS % 2 ? 2 * NANO_STEPS : NANO_STEPS;
        void *pa = a->get(), *pb = b->get();
        loop.run(stream.get(), pa, nanoSteps / period, [&]() {
            void *src = pa, *dst = pb;
            for (unsigned t = 0; t < period; ++t) {
                nanoStep(src, dst, t % NANO_STEPS);
                std::swap(src, dst);
            }
        });
        for (unsigned t = 0; t < nanoSteps % period; ++t) {
            nanoStep(a->get(), b->get(), t % NANO_STEPS);
            a.swap(b);
        }
    }

    void sync() { stream.sync(); }
    const lgdhip_cell_grid& descriptor() const { return g; }
    const GraphLoop& graphs() const { return loop; }

private:
    Stream stream;
    lgdhip_cell_grid g;
    std::unique_ptr<DeviceMemory> a, b, edge;
    int wrap;
    bool uploaded;
    int bx, by;
    GraphLoop loop;

    // one nano step src -> dst: wrapped halo refresh, then the sweep
    void nanoStep(void *src, void *dst, unsigned nano)
    {
        const long rows = g.ny * g.nz;
        const unsigned xblocks = (unsigned)((g.nx + bx - 1) / bx);
        const long groups = (rows + by - 1) / by;
        if ((long)xblocks * groups > 0x7fffffffL) {
            throw std::invalid_argument("generic HIP kernel: grid too large for one launch");
        }
        const int hy = DIM >= 2 ? RADIUS : 0;
        const int hz = DIM == 3 ? RADIUS : 0;
        hipStream_t s = reinterpret_cast<hipStream_t>(stream.get());
        if (wrap) {
            check(lgdhip_cell_grid_fill_halo(src, &g, edge->get(), wrap, stream.get()), "lgdhip_cell_grid_fill_halo");
        }
        GenericHelpers::sweep<CELL, DIM><<<dim3((unsigned)(xblocks * groups)), dim3(bx, by), 0, s>>>(
            static_cast<const CELL *>(src), static_cast<CELL *>(dst), (int)g.nx, rows, (int)g.ny, RADIUS, hy, hz,
            (long)g.px, (long)(g.px * g.py), xblocks, nano);
        GenericHelpers::checkHip(hipGetLastError(), "generic HIP kernel launch");
    }

    void requireBox(const Grid<CELL, Topology>& host) const
    {
        if ((long)host.size() != g.nx * g.ny * g.nz) {
            throw std::invalid_argument("generic HIP kernel: host grid does not match the device grid");
        }
    }
};

}

}

// struct-of-arrays line cells: the SoA family, from the same include
#include <libgeodecomp/hip/soafamily.h>

#endif
