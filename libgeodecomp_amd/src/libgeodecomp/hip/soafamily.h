// HipKernels::SoA<CELL>: the kernel family of struct-of-arrays cells -- a
// cell registered with LIBFLATARRAY_REGISTER_SOA whose API has HasSoA,
// HasFixedCoordsOnlyUpdate and HasUpdateLineX, updated by its own
//     template<class HOOD_OLD, class HOOD_NEW>
//     static void updateLineX(HOOD_OLD& hoodOld, int indexEnd, HOOD_NEW& hoodNew, int nanoStep)
// compiled for gfx950.  This is the sweep the reference leaves empty for
// SoA cells on the GPU (cudasimulator.h:592-594) and runs on the CPU
// through FixedNeighborhoodUpdateFunctor
// (storage/fixedneighborhoodupdatefunctor.h:126-258).
//
// Include this header (or libgeodecomp/hip/genericfamily.h, which includes
// it) from a translation unit compiled by hipcc; HipSimulator<CELL> then
// picks this family for such cells.
//
// Device layout: both ping-pong grids are member-major planes (the
// registration's addressing, member m of element i at
// data + DIM_PROD * offset(m) + i * sizeof(m)) over the logical box grown by
// the stencil radius on every axis that does not wrap (holding the edge
// cell) and not at all on a wrapping axis, like the reference's SoAGrid
// (soagrid.h:578-584).  Rows are padded so the first logical cell of every
// row starts on a 128-byte line for 8-byte members, and the plane size is
// rounded up to 128 elements so every member plane starts aligned.
//
// Sweep: one thread per cell, adjacent lanes on adjacent x (each member
// access of a wave is one contiguous run).  The thread calls
// updateLineX(hoodOld, index + 1, hoodNew, nanoStep) for its own cell: the
// cell's loop runs once; hoodOld is a FixedNeighborhood whose West/East,
// Top/Bottom and South/North offsets wrap the first and last cell of each
// periodic axis exactly as the reference's functor sets them for the
// streak holding that cell.  Whole time steps are replayed as captured
// hipGraphs.
#ifndef LIBGEODECOMP_HIP_SOAFAMILY_H
#define LIBGEODECOMP_HIP_SOAFAMILY_H

#ifndef __HIP__
#error "libgeodecomp/hip/soafamily.h compiles user cells for the GPU: build this translation unit with hipcc"
#endif

#include <hip/hip_runtime.h>

#include <libflatarray/flat_array.hpp>
#include <libgeodecomp/hip/kernelfamilies.h>
#include <libgeodecomp/storage/fixedneighborhood.h>
#include <libgeodecomp/storage/soagrid.h>

#include <climits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace LibGeoDecomp {

namespace HipKernels {

namespace SoAHelpers {

// extents and strides of one device SoA grid, passed by value to kernels
struct Geometry
{
    int nx, ny, nz;     // logical extents (1 past the grid's dimensionality)
    int ox, oy, oz;     // element offset of logical (0, 0, 0)
    long dimX, dimXY;   // row and plane pitch in elements
    long prod;          // elements per member plane
};

template<typename CELL, bool WX, bool WY, bool WZ>
__global__ void __launch_bounds__(256) sweep(char *src, char *dst, Geometry geo, unsigned xblocks, unsigned nanoStep)
{
    const unsigned xb = blockIdx.x % xblocks;
    const long row = (long)(blockIdx.x / xblocks) * blockDim.y + threadIdx.y;
    const int x = (int)(xb * blockDim.x + threadIdx.x);
    if (x >= geo.nx || row >= (long)geo.ny * geo.nz) {
        return;
    }
    const int y = (int)(row % geo.ny);
    const int z = (int)(row / geo.ny);
    const long index = (z + geo.oz) * geo.dimXY + (y + geo.oy) * geo.dimX + (x + geo.ox);
    LibFlatArray::soa_accessor<CELL> accOld(src, index, geo.dimX, geo.dimXY, geo.prod);
    LibFlatArray::soa_accessor<CELL> hoodNew(dst, index, geo.dimX, geo.dimXY, geo.prod);
    // the wrap offsets of the streak functor for the streak of one cell
    const long west = (WX && x == 0) ? geo.nx : 0;
    const long east = (WX && x == geo.nx - 1) ? -geo.nx : 0;
    const long top = (WY && y == 0) ? geo.ny : 0;
    const long bottom = (WY && y == geo.ny - 1) ? -geo.ny : 0;
    const long south = (WZ && z == 0) ? geo.nz : 0;
    const long north = (WZ && z == geo.nz - 1) ? -geo.nz : 0;
    FixedNeighborhood<CELL> hoodOld(accOld, west, east, top, bottom, south, north);
    long indexEnd = index + 1;
    CELL::updateLineX(hoodOld, indexEnd, hoodNew, nanoStep);
}

inline void checkHip(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        throw std::runtime_error(std::string(what) + " failed: " + hipGetErrorString(e));
    }
}

}

template<typename CELL>
class SoA
{
public:
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    typedef typename APITraits::SelectStencil<CELL>::Value Stencil;
    typedef LibFlatArray::soa_traits<CELL> Traits;
    typedef LibFlatArray::soa_accessor<CELL> Accessor;
    static const int DIM = Topology::DIM;
    static const int RADIUS = Stencil::RADIUS;
    static const unsigned NANO_STEPS = APITraits::SelectNanoSteps<CELL>::VALUE;
    static const bool WX = Topology::template WrapsAxis<0>::VALUE;
    static const bool WY = DIM > 1 && Topology::template WrapsAxis<1>::VALUE;
    static const bool WZ = DIM > 2 && Topology::template WrapsAxis<2>::VALUE;

    static_assert(APITraits::SelectSoAStreakUpdate<CELL>::VALUE,
                  "the SoA HIP kernel family runs cells with HasSoA, HasFixedCoordsOnlyUpdate and HasUpdateLineX");
    static_assert(RADIUS >= 1 && RADIUS <= 16, "stencil radius must be in [1, 16]");
    static_assert(RADIUS == 1 || !(WX || WY || WZ),
                  "struct-of-arrays cells with a wrapped axis need a stencil radius of 1: the neighbourhood wraps "
                  "the first and last cell of an axis (storage/fixedneighborhood.h)");

    // tuning.tile.bx / by: work-group shape (x cells, rows); zero = 64 x 4
    // (1D: 256 x 1)
    SoA(const Coord<DIM>& dim, const Tuning& tuning) :
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
        if (bx * by > 256 || bx % 64 != 0) {
            throw std::invalid_argument("SoA HIP kernel: work-group width must be a multiple of 64 and width x "
                                        "rows at most 256");
        }
        Coord<3> radii;
        for (int d = 0; d < 3; ++d) {
            radii[d] = (d < DIM && !Topology::wrapsAxis(d)) ? RADIUS : 0;
        }
        layout = SoALayout(SoAGridHelpers::expand(dim, 1), radii, 16, 128);
        // user code may hold indices in an int (updateLineX(..., int indexEnd, ...))
        if (layout.prod > (long)INT_MAX) {
            throw std::invalid_argument("SoA HIP kernel: grid has more than 2^31 elements per member plane");
        }
        geo.nx = layout.logical.x();
        geo.ny = layout.logical.y();
        geo.nz = layout.logical.z();
        geo.ox = layout.origin.x();
        geo.oy = layout.origin.y();
        geo.oz = layout.origin.z();
        geo.dimX = layout.dimX;
        geo.dimXY = layout.dimX * layout.dimY;
        geo.prod = layout.prod;
        bytes = (std::size_t)layout.prod * Traits::BYTES_PER_CELL;
        a.reset(new DeviceMemory(bytes));
        b.reset(new DeviceMemory(bytes));
    }

    // whole grid: the logical cells and the edge cell in every padding
    // element, converted to member-major planes on the host, one copy
    void upload(const Grid<CELL, Topology>& host)
    {
        requireBox(host);
        std::vector<char> staging(bytes);
        const CELL& edge = host.getEdge();
        for (long Z = 0; Z < layout.dimZ; ++Z) {
            for (long Y = 0; Y < layout.dimY; ++Y) {
                for (long X = 0; X < layout.dimX; ++X) {
                    Accessor acc(staging.data(), (Z * layout.dimY + Y) * layout.dimX + X, geo.dimX, geo.dimXY,
                                 geo.prod);
                    if (layout.inside(X, Y, Z)) {
                        const long x = X - geo.ox, y = Y - geo.oy, z = Z - geo.oz;
                        acc << host.data()[(z * geo.ny + y) * geo.nx + x];
                    } else {
                        acc << edge;
                    }
                }
            }
        }
        check(lgdhip_memcpy_h2d(a->get(), staging.data(), bytes, stream.get()), "lgdhip_memcpy_h2d");
        if (!uploaded) {
            // both grids start from the initial state, as the reference's
            // simulators initialise both
            check(lgdhip_memcpy_h2d(b->get(), staging.data(), bytes, stream.get()), "lgdhip_memcpy_h2d");
            uploaded = true;
        }
    }

    // the region's cells only: per streak and member plane one contiguous copy
    void uploadRegion(const Grid<CELL, Topology>& host, const Region<DIM>& region)
    {
        requireBox(host);
        const std::vector<Streak<DIM> > streaks = relativeStreaks(region, host.boundingBox());
        std::vector<char> line;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<DIM>& st = streaks[k];
            const long n = st.endX - st.origin.x();
            const long y = coordOf(st.origin, 1), z = coordOf(st.origin, 2);
            const CELL *cells = host.data() + (z * geo.ny + y) * geo.nx + st.origin.x();
            // the streak as a member-major block of n elements
            line.assign((std::size_t)n * Traits::BYTES_PER_CELL, 0);
            for (long i = 0; i < n; ++i) {
                Accessor acc(line.data(), i, n, n, n);
                acc << cells[i];
            }
            const long first = layout.index(st.origin.x(), y, z);
            Traits::each_member([&](std::size_t offset, std::size_t size, std::size_t count) {
                for (std::size_t e = 0; e < count; ++e) {
                    const std::size_t plane = offset + e * size;
                    char *dst = static_cast<char *>(a->get()) + (std::size_t)geo.prod * plane + first * size;
                    check(lgdhip_memcpy_h2d(dst, line.data() + (std::size_t)n * plane, (std::size_t)n * size,
                                            stream.get()),
                          "lgdhip_memcpy_h2d");
                }
            });
        }
    }

    void download(Grid<CELL, Topology> *host)
    {
        requireBox(*host);
        std::vector<char> staging(bytes);
        check(lgdhip_memcpy_d2h(staging.data(), a->get(), bytes, stream.get()), "lgdhip_memcpy_d2h");
        for (long z = 0; z < geo.nz; ++z) {
            for (long y = 0; y < geo.ny; ++y) {
                for (long x = 0; x < geo.nx; ++x) {
                    Accessor acc(staging.data(), layout.index(x, y, z), geo.dimX, geo.dimXY, geo.prod);
                    acc >> host->data()[(z * geo.ny + y) * geo.nx + x];
                }
            }
        }
    }

    // the region's cells only: per streak one strided copy of its segment
    // of every member plane, converted back to cells on the host
    void downloadRegion(Grid<CELL, Topology> *host, const Region<DIM>& region)
    {
        requireBox(*host);
        const std::vector<Streak<DIM> > streaks = relativeStreaks(region, host->boundingBox());
        std::vector<char> line;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<DIM>& st = streaks[k];
            const long n = st.endX - st.origin.x();
            const long y = coordOf(st.origin, 1), z = coordOf(st.origin, 2);
            const long first = layout.index(st.origin.x(), y, z);
            line.assign((std::size_t)n * Traits::BYTES_PER_CELL, 0);
            Traits::each_member([&](std::size_t offset, std::size_t size, std::size_t count) {
                // the `count` planes of this member lie geo.prod * size bytes apart
                const std::size_t plane = offset;
                const char *src = static_cast<const char *>(a->get()) + (std::size_t)geo.prod * plane + first * size;
                check(lgdhip_memcpy2d_d2h(line.data() + (std::size_t)n * plane, (std::size_t)n * size, src,
                                          (std::size_t)geo.prod * size, (std::size_t)n * size, count, stream.get()),
                      "lgdhip_memcpy2d_d2h");
            });
            CELL *cells = host->data() + (z * geo.ny + y) * geo.nx + st.origin.x();
            for (long i = 0; i < n; ++i) {
                Accessor acc(line.data(), i, n, n, n);
                acc >> cells[i];
            }
        }
    }

    // nanoSteps is a multiple of NANO_STEPS (whole steps); a period of
    // lcm(2, NANO_STEPS) nano steps is replayed as one captured graph
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
    const SoALayout& soaLayout() const { return layout; }
    const GraphLoop& graphs() const { return loop; }

private:
    Stream stream;
    SoALayout layout;
    SoAHelpers::Geometry geo;
    std::size_t bytes;
    std::unique_ptr<DeviceMemory> a, b;
    bool uploaded;
    int bx, by;
    GraphLoop loop;

    void nanoStep(void *src, void *dst, unsigned nano)
    {
        const long rows = (long)geo.ny * geo.nz;
        const unsigned xblocks = (unsigned)((geo.nx + bx - 1) / bx);
        const long groups = (rows + by - 1) / by;
        if ((long)xblocks * groups > 0x7fffffffL) {
            throw std::invalid_argument("SoA HIP kernel: grid too large for one launch");
        }
        hipStream_t s = reinterpret_cast<hipStream_t>(stream.get());
        SoAHelpers::sweep<CELL, WX, WY, WZ><<<dim3((unsigned)(xblocks * groups)), dim3(bx, by), 0, s>>>(
            static_cast<char *>(src), static_cast<char *>(dst), geo, xblocks, nano);
        SoAHelpers::checkHip(hipGetLastError(), "SoA HIP kernel launch");
    }

    void requireBox(const Grid<CELL, Topology>& host) const
    {
        if ((long)host.size() != (long)geo.nx * geo.ny * geo.nz) {
            throw std::invalid_argument("SoA HIP kernel: host grid does not match the device grid");
        }
    }
};

}

}

#endif
