// Kernel families: host-side adapters between a user Cell stored in the
// host Grid and the hand-written gfx950 sweeps behind the lgdhip C-ABI.
//
// A family owns the device-resident state of one simulation (padded SoA
// grid, ping-pong buffers, workspaces, a HIP stream) and offers
//     upload(hostGrid)   host cells -> device state
//     download(&grid)    device state -> host cells
//     run(nanoSteps)     advance the device state
// A cell opts in with APITraits::HasHipKernel<FAMILY> or, after its
// definition, LIBGEODECOMP_REGISTER_HIP_KERNEL(CELL, FAMILY).  The family's
// kernels compute exactly the update() of the cell they stand for; the
// contract is checked by running the same cell through the CPU oracle.
//
// Families:
//   Jacobi3DF64<CELL, double, &CELL::member>
//       JacobiCellFixedHood::update (src/testbed/performancetests/main.cpp:1194-1205)
//   Life2D<CELL, bool|char|uint8, &CELL::member>
//       ConwayCell::update (src/examples/gameoflife/main.cpp:36-59)
//   LBMD3Q19F32<CELL>  (CELL has float f[19], unsigned char flags and the
//       static parameters CELL::omega, CELL::lidVelocity[3])
//       D3Q19 BGK pull cell after src/examples/latticeboltzmann/main.cpp:92-229
#ifndef LIBGEODECOMP_HIP_KERNELFAMILIES_H
#define LIBGEODECOMP_HIP_KERNELFAMILIES_H

#include <libgeodecomp/geometry/stencils.h>
#include <libgeodecomp/misc/apitraits.h>
#include <libgeodecomp/storage/grid.h>

#include <lgdhip.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace LibGeoDecomp {

namespace HipKernels {

// Converts a non-zero lgdhip status into std::runtime_error (the
// reference's CUDAUtil::checkForError throws the same type,
// src/libgeodecomp/misc/cudautil.h:48-55).
inline void check(int rc, const char *what)
{
    if (rc != 0) {
        const char *detail = lgdhip_last_error();
        throw std::runtime_error(std::string(what) + " failed: " +
                                 ((detail && *detail) ? detail : lgdhip_error_string(rc)));
    }
}

class DeviceMemory
{
public:
    DeviceMemory() : ptr(0), bytes(0) {}
    explicit DeviceMemory(std::size_t n) : ptr(0), bytes(n) { check(lgdhip_malloc(&ptr, n), "lgdhip_malloc"); }
    ~DeviceMemory() { lgdhip_free(ptr); }
    DeviceMemory(const DeviceMemory&) = delete;
    DeviceMemory& operator=(const DeviceMemory&) = delete;

    void *get() const { return ptr; }
    std::size_t size() const { return bytes; }

private:
    void *ptr;
    std::size_t bytes;
};

class Stream
{
public:
    Stream() : s(0) { check(lgdhip_stream_create(&s), "lgdhip_stream_create"); }
    ~Stream() { lgdhip_stream_destroy(s); }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;

    lgdhip_stream_t get() const { return s; }
    void sync() const { check(lgdhip_stream_synchronize(s), "lgdhip_stream_synchronize"); }

private:
    lgdhip_stream_t s;
};

// Tuning knobs handed down by HipSimulator (blockSize -> bx/by/bz).
struct Tuning
{
    lgdhip_tile tile;
    bool useTile;

    Tuning() : useTile(false) { std::memset(&tile, 0, sizeof(tile)); }
    const lgdhip_tile *get() const { return useTile ? &tile : 0; }
};

// wrapped axes of the cell's topology as a mask (bit d = axis d)
template<typename CELL>
int wrapMask()
{
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    int mask = 0;
    for (int d = 0; d < Topology::DIM; ++d) {
        if (Topology::wrapsAxis(d)) {
            mask |= 1 << d;
        }
    }
    return mask;
}

template<typename CELL>
void requireSingleNanoStep()
{
    static_assert(APITraits::SelectNanoSteps<CELL>::VALUE == 1, "kernel family supports one nano step per step");
}

template<typename CELL>
void requireCubeSingleNanoStep()
{
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    for (int d = 0; d < Topology::DIM; ++d) {
        if (Topology::wrapsAxis(d)) {
            throw std::logic_error("this HIP kernel family implements the Cube topology only");
        }
    }
    static_assert(APITraits::SelectNanoSteps<CELL>::VALUE == 1, "kernel family supports one nano step per step");
}

// ---------------------------------------------------------------------
// Jacobi3D FP64 7-point
// ---------------------------------------------------------------------
template<typename CELL, typename VALUE, VALUE CELL::*MEMBER>
class Jacobi3DF64
{
public:
    static_assert(std::is_same<VALUE, double>::value, "Jacobi3DF64 needs a double member");
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    static const int DIM = 3;
    static_assert(Topology::DIM == 3, "Jacobi3DF64 is a 3D family");

    // Cube: the tuned (temporally blocked) sweeps.  Wrapped axes (Torus
    // topologies): single sweeps, each after a periodic halo refresh.
    Jacobi3DF64(const Coord<3>& dim, const Tuning& tuning) :
        tuning(tuning),
        wrap(wrapMask<CELL>()),
        edge(0.0)
    {
        requireSingleNanoStep<CELL>();
        check(lgdhip_grid_make(&g, 3, dim.x(), dim.y(), dim.z(), 1, 8, 1), "lgdhip_grid_make");
        const std::size_t n = (std::size_t)lgdhip_grid_bytes(&g);
        a.reset(new DeviceMemory(n));
        b.reset(new DeviceMemory(n));
    }

    void upload(const Grid<CELL, Topology>& host)
    {
        std::vector<double> dense(host.size());
        for (std::size_t i = 0; i < dense.size(); ++i) {
            dense[i] = host.data()[i].*MEMBER;
        }
        edge = host.getEdge().*MEMBER;
        check(lgdhip_grid_fill_edge(a->get(), &g, &edge, stream.get()), "lgdhip_grid_fill_edge");
        check(lgdhip_grid_fill_edge(b->get(), &g, &edge, stream.get()), "lgdhip_grid_fill_edge");
        check(lgdhip_grid_load(a->get(), &g, dense.data(), stream.get()), "lgdhip_grid_load");
    }

    void download(Grid<CELL, Topology> *host)
    {
        std::vector<double> dense(host->size());
        check(lgdhip_grid_store(a->get(), &g, dense.data(), stream.get()), "lgdhip_grid_store");
        for (std::size_t i = 0; i < dense.size(); ++i) {
            host->data()[i].*MEMBER = dense[i];
        }
    }

    void run(unsigned nanoSteps)
    {
        if (wrap) {
            lgdhip_tile single;
            std::memset(&single, 0, sizeof(single));
            single.variant = 2;  // a single sweep that reads its halo from memory
            for (unsigned t = 0; t < nanoSteps; ++t) {
                int inB = 0;
                check(lgdhip_grid_fill_halo(a->get(), &g, &edge, wrap, stream.get()), "lgdhip_grid_fill_halo");
                check(lgdhip_jacobi3d_f64_run((double *)a->get(), (double *)b->get(), &g, 1, &single, &inB,
                                              stream.get()),
                      "lgdhip_jacobi3d_f64_run");
                a.swap(b);
            }
            return;
        }
        int inB = 0;
        check(lgdhip_jacobi3d_f64_run((double *)a->get(), (double *)b->get(), &g, (int)nanoSteps, tuning.get(), &inB,
                                      stream.get()),
              "lgdhip_jacobi3d_f64_run");
        if (inB) {
            a.swap(b);
        }
    }

    void sync() { stream.sync(); }
    const lgdhip_grid& descriptor() const { return g; }

private:
    Tuning tuning;
    Stream stream;
    lgdhip_grid g;
    std::unique_ptr<DeviceMemory> a, b;
    int wrap;
    double edge;
};

// ---------------------------------------------------------------------
// Game of Life (uint8 device grid, bit-packed temporal-blocking sweeps)
// ---------------------------------------------------------------------
template<typename CELL, typename VALUE, VALUE CELL::*MEMBER>
class Life2D
{
public:
    static_assert(sizeof(VALUE) == 1, "Life2D needs a one-byte alive member (bool, char, uint8_t)");
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    static const int DIM = 2;
    static_assert(Topology::DIM == 2, "Life2D is a 2D family");

    // Cube: bit-packed temporally blocked sweeps.  Wrapped axes (Torus
    // topologies): byte-grid single generations, each after a periodic halo
    // refresh (the state then stays in the byte grids).
    Life2D(const Coord<2>& dim, const Tuning& tuning) :
        tuning(tuning),
        edgeAlive(0),
        wrap(wrapMask<CELL>())
    {
        requireSingleNanoStep<CELL>();
        check(lgdhip_grid_make(&g, 2, dim.x(), dim.y(), 1, 1, 1, 1), "lgdhip_grid_make");
        grid.reset(new DeviceMemory((std::size_t)lgdhip_grid_bytes(&g)));
        if (wrap) {
            grid2.reset(new DeviceMemory((std::size_t)lgdhip_grid_bytes(&g)));
        }
        const std::size_t ws = lgdhip_life2d_workspace_bytes(&g);
        packed.reset(new DeviceMemory(ws));
        pa = (uint32_t *)packed->get();
        pb = (uint32_t *)((char *)packed->get() + ws / 2);
    }

    void upload(const Grid<CELL, Topology>& host)
    {
        std::vector<uint8_t> dense(host.size());
        for (std::size_t i = 0; i < dense.size(); ++i) {
            dense[i] = (host.data()[i].*MEMBER) ? 1 : 0;
        }
        edgeAlive = (host.getEdge().*MEMBER) ? 1 : 0;
        const uint8_t edge = (uint8_t)edgeAlive;
        check(lgdhip_grid_fill_edge(grid->get(), &g, &edge, stream.get()), "lgdhip_grid_fill_edge");
        check(lgdhip_grid_load(grid->get(), &g, dense.data(), stream.get()), "lgdhip_grid_load");
        if (wrap) {
            check(lgdhip_grid_fill_edge(grid2->get(), &g, &edge, stream.get()), "lgdhip_grid_fill_edge");
            return;
        }
        check(lgdhip_life2d_pack((const uint8_t *)grid->get(), &g, pa, stream.get()), "lgdhip_life2d_pack");
    }

    void download(Grid<CELL, Topology> *host)
    {
        if (!wrap) {
            check(lgdhip_life2d_unpack(pa, &g, (uint8_t *)grid->get(), stream.get()), "lgdhip_life2d_unpack");
        }
        std::vector<uint8_t> dense(host->size());
        check(lgdhip_grid_store(grid->get(), &g, dense.data(), stream.get()), "lgdhip_grid_store");
        for (std::size_t i = 0; i < dense.size(); ++i) {
            host->data()[i].*MEMBER = static_cast<VALUE>(dense[i]);
        }
    }

    void run(unsigned nanoSteps)
    {
        if (wrap) {
            const uint8_t edge = (uint8_t)edgeAlive;
            for (unsigned t = 0; t < nanoSteps; ++t) {
                int inB = 0;
                check(lgdhip_grid_fill_halo(grid->get(), &g, &edge, wrap, stream.get()), "lgdhip_grid_fill_halo");
                check(lgdhip_life2d_u8_run((uint8_t *)grid->get(), (uint8_t *)grid2->get(), &g, 1, &inB,
                                           stream.get()),
                      "lgdhip_life2d_u8_run");
                grid.swap(grid2);
            }
            return;
        }
        int inB = 0;
        check(lgdhip_life2d_packed_run(pa, pb, &g, edgeAlive, (int)nanoSteps, tuning.get(), &inB, stream.get()),
              "lgdhip_life2d_packed_run");
        if (inB) {
            std::swap(pa, pb);
        }
    }

    void sync() { stream.sync(); }

private:
    Tuning tuning;
    Stream stream;
    lgdhip_grid g;
    std::unique_ptr<DeviceMemory> grid, grid2, packed;
    uint32_t *pa, *pb;
    int edgeAlive;
    int wrap;
};

// ---------------------------------------------------------------------
// D3Q19 lid-driven cavity, FP32 populations
// ---------------------------------------------------------------------
template<typename CELL>
class LBMD3Q19F32
{
public:
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    static const int DIM = 3;
    static_assert(Topology::DIM == 3, "LBMD3Q19F32 is a 3D family");

    LBMD3Q19F32(const Coord<3>& dim, const Tuning& tuning) :
        tuning(tuning)
    {
        requireCubeSingleNanoStep<CELL>();
        check(lgdhip_grid_make(&g, 3, dim.x(), dim.y(), dim.z(), 1, 4, 19), "lgdhip_grid_make");
        // the 19 populations of each z slice stored together: fewer
        // concurrent HBM streams per sweep than whole member planes
        check(lgdhip_grid_set_layout(&g, LGDHIP_LAYOUT_Z_SLABS), "lgdhip_grid_set_layout");
        check(lgdhip_grid_make(&fg, 3, dim.x(), dim.y(), dim.z(), 1, 1, 1), "lgdhip_grid_make");
        a.reset(new DeviceMemory((std::size_t)lgdhip_grid_bytes(&g)));
        b.reset(new DeviceMemory((std::size_t)lgdhip_grid_bytes(&g)));
        flags.reset(new DeviceMemory((std::size_t)lgdhip_grid_bytes(&fg)));
    }

    // lattice weight of population q (rest, axis, diagonal)
    static double weight(int q) { return q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }

    // the device keeps the shifted populations g = f - w (see lgdhip.h)
    void upload(const Grid<CELL, Topology>& host)
    {
        const std::size_t n = host.size();
        std::vector<float> dense(19 * n);
        std::vector<uint8_t> fl(n);
        for (std::size_t i = 0; i < n; ++i) {
            const CELL& c = host.data()[i];
            for (int q = 0; q < 19; ++q) {
                dense[q * n + i] = (float)((double)c.f[q] - weight(q));
            }
            fl[i] = c.flags;
        }
        float edge[19];
        for (int q = 0; q < 19; ++q) {
            edge[q] = (float)((double)host.getEdge().f[q] - weight(q));
        }
        const uint8_t edgeFlags = host.getEdge().flags;
        check(lgdhip_grid_fill_edge(a->get(), &g, edge, stream.get()), "lgdhip_grid_fill_edge");
        check(lgdhip_grid_fill_edge(b->get(), &g, edge, stream.get()), "lgdhip_grid_fill_edge");
        check(lgdhip_grid_fill_edge(flags->get(), &fg, &edgeFlags, stream.get()), "lgdhip_grid_fill_edge");
        check(lgdhip_grid_load(a->get(), &g, dense.data(), stream.get()), "lgdhip_grid_load");
        check(lgdhip_grid_load(flags->get(), &fg, fl.data(), stream.get()), "lgdhip_grid_load");
    }

    void download(Grid<CELL, Topology> *host)
    {
        const std::size_t n = host->size();
        std::vector<float> dense(19 * n);
        check(lgdhip_grid_store(a->get(), &g, dense.data(), stream.get()), "lgdhip_grid_store");
        for (std::size_t i = 0; i < n; ++i) {
            for (int q = 0; q < 19; ++q) {
                host->data()[i].f[q] = (float)((double)dense[q * n + i] + weight(q));
            }
        }
    }

    void run(unsigned nanoSteps)
    {
        lgdhip_lbm_params p;
        p.omega = (float)CELL::omega;
        p.lid_ux = (float)CELL::lidVelocity[0];
        p.lid_uy = (float)CELL::lidVelocity[1];
        p.lid_uz = (float)CELL::lidVelocity[2];
        int inB = 0;
        check(lgdhip_lbm_d3q19_f32_run((float *)a->get(), (float *)b->get(), (const uint8_t *)flags->get(), &g, &p,
                                       (int)nanoSteps, tuning.get(), &inB, stream.get()),
              "lgdhip_lbm_d3q19_f32_run");
        if (inB) {
            a.swap(b);
        }
    }

    void sync() { stream.sync(); }

private:
    Tuning tuning;
    Stream stream;
    lgdhip_grid g, fg;
    std::unique_ptr<DeviceMemory> a, b, flags;
};

}

}

#endif
