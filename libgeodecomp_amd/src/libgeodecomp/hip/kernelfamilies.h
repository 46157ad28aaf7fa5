// Kernel families: host-side adapters between a user Cell stored in the
// host Grid and the hand-written gfx950 sweeps behind the lgdhip C-ABI.
//
// A family owns the device-resident state of one simulation (padded SoA
// grid, ping-pong buffers, workspaces, a HIP stream) and offers
//     upload(hostGrid)                  host cells -> device state
//     uploadRegion(hostGrid, region)    only the region's cells (optional;
//                                       HipSimulator uses it after steerers)
//     download(&grid)                   device state -> host cells
//     downloadRegion(&grid, region)     only the region's cells (optional;
//                                       HipSimulator uses it for the rows
//                                       its Writers read)
//     run(nanoSteps)                    advance the device state
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

#include <libgeodecomp/geometry/region.h>
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

// A fixed sequence of launches (one "period" of a family's time loop, e.g.
// two nano steps that return the ping-pong buffers to their places)
// captured once into a hipGraph per buffer arrangement and replayed: the
// multi-launch steps of the time loop go to the GPU as one graph launch per
// period instead of one launch call per kernel (the reference launches
// every nano step separately, cudasimulator.h:467-470).  Only kernel
// launches are captured (no memset or copy nodes).
class GraphLoop
{
public:
    GraphLoop() { clear(); }
    ~GraphLoop() { reset(); }
    GraphLoop(const GraphLoop&) = delete;
    GraphLoop& operator=(const GraphLoop&) = delete;

    // enqueue() `count` times on `stream`; key names the arrangement the
    // launches of one period depend on (e.g. the current buffer's address)
    template<typename ENQUEUE>
    void run(lgdhip_stream_t stream, const void *key, unsigned long long count, ENQUEUE enqueue)
    {
        if (count < 2) {
            for (unsigned long long i = 0; i < count; ++i) {
                enqueue();
            }
            return;
        }
        lgdhip_graph_t graph = find(key);
        if (!graph) {
            check(lgdhip_graph_begin(stream), "lgdhip_graph_begin");
            try {
                enqueue();
            } catch (...) {
                lgdhip_graph_t broken = 0;
                if (lgdhip_graph_end(stream, &broken) == 0) {
                    lgdhip_graph_destroy(broken);
                }
                throw;
            }
            check(lgdhip_graph_end(stream, &graph), "lgdhip_graph_end");
            store(key, graph);
            ++captures;
        }
        for (unsigned long long i = 0; i < count; ++i) {
            check(lgdhip_graph_launch(graph, stream), "lgdhip_graph_launch");
        }
        replays += count;
    }

    // destroys the captured graphs (after the buffers they name changed)
    void reset()
    {
        for (int i = 0; i < SLOTS; ++i) {
            if (graphs[i]) {
                lgdhip_graph_destroy(graphs[i]);
            }
        }
        clear();
    }

    unsigned long long graphsCaptured() const { return captures; }
    unsigned long long graphLaunches() const { return replays; }

private:
    static const int SLOTS = 2;
    const void *keys[SLOTS];
    lgdhip_graph_t graphs[SLOTS];
    unsigned long long captures = 0, replays = 0;

    void clear()
    {
        for (int i = 0; i < SLOTS; ++i) {
            keys[i] = 0;
            graphs[i] = 0;
        }
    }

    lgdhip_graph_t find(const void *key) const
    {
        for (int i = 0; i < SLOTS; ++i) {
            if (graphs[i] && keys[i] == key) {
                return graphs[i];
            }
        }
        return 0;
    }

    void store(const void *key, lgdhip_graph_t graph)
    {
        for (int i = 0; i < SLOTS; ++i) {
            if (!graphs[i]) {
                keys[i] = key;
                graphs[i] = graph;
                return;
            }
        }
        lgdhip_graph_destroy(graphs[0]);
        keys[0] = keys[1];
        graphs[0] = graphs[1];
        keys[1] = key;
        graphs[1] = graph;
    }
};

// Tuning knobs handed down by HipSimulator (blockSize -> bx/by/bz).
struct Tuning
{
    lgdhip_tile tile;
    bool useTile;

    Tuning() : useTile(false) { std::memset(&tile, 0, sizeof(tile)); }
    const lgdhip_tile *get() const { return useTile ? &tile : 0; }
};

// element index of (x, y, z) of member m in a padded grid, in either member
// layout (the address formula documented with lgdhip_grid)
inline int64_t elementIndex(const lgdhip_grid& g, int m, int64_t x, int64_t y, int64_t z)
{
    const bool slabs = g.layout == LGDHIP_LAYOUT_Z_SLABS;
    const int64_t ms = slabs ? g.sx * g.sy : g.sx * g.sy * g.sz;
    const int64_t zs = slabs ? (int64_t)g.members * g.sx * g.sy : g.sx * g.sy;
    return m * ms + (z + g.oz) * zs + (y + g.oy) * g.sx + (x + g.ox);
}

// the region's streaks relative to the host grid's origin
template<int DIM>
std::vector<Streak<DIM> > relativeStreaks(const Region<DIM>& region, const CoordBox<DIM>& box)
{
    std::vector<Streak<DIM> > out = region.streaks();
    for (std::size_t i = 0; i < out.size(); ++i) {
        out[i].endX -= box.origin.x();
        out[i].origin = out[i].origin - box.origin;
    }
    return out;
}

template<int DIM>
int64_t coordOf(const Coord<DIM>& c, int d)
{
    return d < DIM ? c[d] : 0;
}

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

    // The tuned (temporally blocked) sweeps; wrapped axes (Torus topologies)
    // take their periodic neighbours inside the same blocked kernels.
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

    void uploadRegion(const Grid<CELL, Topology>& host, const Region<3>& region)
    {
        std::vector<double> line;
        const std::vector<Streak<3> > streaks = relativeStreaks(region, host.boundingBox());
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<3>& st = streaks[k];
            line.resize(st.endX - st.origin.x());
            for (std::size_t i = 0; i < line.size(); ++i) {
                line[i] = host.data()[(st.origin + Coord<3>((int)i, 0, 0)).toIndex(host.boundingBox().dimensions)].*MEMBER;
            }
            double *dst = (double *)a->get() + elementIndex(g, 0, st.origin.x(), st.origin.y(), st.origin.z());
            check(lgdhip_memcpy_h2d(dst, line.data(), line.size() * sizeof(double), stream.get()), "lgdhip_memcpy_h2d");
        }
    }

    void download(Grid<CELL, Topology> *host)
    {
        std::vector<double> dense(host->size());
        check(lgdhip_grid_store(a->get(), &g, dense.data(), stream.get()), "lgdhip_grid_store");
        for (std::size_t i = 0; i < dense.size(); ++i) {
            host->data()[i].*MEMBER = dense[i];
        }
    }

    // one contiguous copy per streak
    void downloadRegion(Grid<CELL, Topology> *host, const Region<3>& region)
    {
        const Coord<3> dim = host->boundingBox().dimensions;
        const std::vector<Streak<3> > streaks = relativeStreaks(region, host->boundingBox());
        std::vector<double> line;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<3>& st = streaks[k];
            line.resize(st.endX - st.origin.x());
            const double *src = (const double *)a->get() + elementIndex(g, 0, st.origin.x(), st.origin.y(),
                                                                           st.origin.z());
            check(lgdhip_memcpy_d2h(line.data(), src, line.size() * sizeof(double), stream.get()),
                  "lgdhip_memcpy_d2h");
            CELL *cells = host->data() + st.origin.toIndex(dim);
            for (std::size_t i = 0; i < line.size(); ++i) {
                cells[i].*MEMBER = line[i];
            }
        }
    }

    void run(unsigned nanoSteps)
    {
        int inB = 0;
        if (wrap) {
            check(lgdhip_jacobi3d_f64_run_periodic((double *)a->get(), (double *)b->get(), &g, (int)nanoSteps, wrap,
                                                   edge, tuning.get(), &inB, stream.get()),
                  "lgdhip_jacobi3d_f64_run_periodic");
        } else {
            check(lgdhip_jacobi3d_f64_run((double *)a->get(), (double *)b->get(), &g, (int)nanoSteps, tuning.get(),
                                          &inB, stream.get()),
                  "lgdhip_jacobi3d_f64_run");
        }
        if (inB) {
            a.swap(b);
        }
    }

    void sync() { stream.sync(); }
    const lgdhip_grid& descriptor() const { return g; }
    const GraphLoop& graphs() const { return loop; }

private:
    Tuning tuning;
    Stream stream;
    lgdhip_grid g;
    std::unique_ptr<DeviceMemory> a, b;
    int wrap;
    double edge;
    GraphLoop loop;
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

    // Bit-packed temporally blocked sweeps; wrapped axes (Torus topologies)
    // load their periodic images inside the same kernels.  Only a wrapped x
    // axis whose length is not a multiple of 32 (the packed row would wrap
    // inside a word) runs byte-grid single generations, each after a
    // periodic halo refresh (the state then stays in the byte grids).
    Life2D(const Coord<2>& dim, const Tuning& tuning) :
        tuning(tuning),
        edgeAlive(0),
        wrap(wrapMask<CELL>()),
        bytes((wrap & 1) && dim.x() % 32 != 0)
    {
        requireSingleNanoStep<CELL>();
        check(lgdhip_grid_make(&g, 2, dim.x(), dim.y(), 1, 1, 1, 1), "lgdhip_grid_make");
        grid.reset(new DeviceMemory((std::size_t)lgdhip_grid_bytes(&g)));
        if (bytes) {
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
        if (bytes) {
            check(lgdhip_grid_fill_edge(grid2->get(), &g, &edge, stream.get()), "lgdhip_grid_fill_edge");
            return;
        }
        check(lgdhip_life2d_pack((const uint8_t *)grid->get(), &g, pa, stream.get()), "lgdhip_life2d_pack");
    }

    // byte-grid mode: bytes into the byte grid; otherwise the packed words
    // covering each streak, packed from the (current) host row
    void uploadRegion(const Grid<CELL, Topology>& host, const Region<2>& region)
    {
        const Coord<2> dim = host.boundingBox().dimensions;
        const std::vector<Streak<2> > streaks = relativeStreaks(region, host.boundingBox());
        const int64_t nw = (g.nx + 31) / 32;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<2>& st = streaks[k];
            const int y = st.origin.y();
            const CELL *row = host.data() + (std::size_t)y * dim.x();
            if (bytes) {
                std::vector<uint8_t> line(st.endX - st.origin.x());
                for (std::size_t i = 0; i < line.size(); ++i) {
                    line[i] = (row[st.origin.x() + i].*MEMBER) ? 1 : 0;
                }
                uint8_t *dst = (uint8_t *)grid->get() + elementIndex(g, 0, st.origin.x(), y, 0);
                check(lgdhip_memcpy_h2d(dst, line.data(), line.size(), stream.get()), "lgdhip_memcpy_h2d");
                continue;
            }
            const int64_t w0 = st.origin.x() / 32, w1 = (st.endX - 1) / 32;
            std::vector<uint32_t> words(w1 - w0 + 1, 0u);
            for (int64_t w = w0; w <= w1; ++w) {
                for (int i = 0; i < 32 && 32 * w + i < g.nx; ++i) {
                    if (row[32 * w + i].*MEMBER) {
                        words[w - w0] |= 1u << i;
                    }
                }
            }
            check(lgdhip_memcpy_h2d(pa + y * nw + w0, words.data(), words.size() * 4, stream.get()),
                  "lgdhip_memcpy_h2d");
        }
    }

    void download(Grid<CELL, Topology> *host)
    {
        if (!bytes) {
            check(lgdhip_life2d_unpack(pa, &g, (uint8_t *)grid->get(), stream.get()), "lgdhip_life2d_unpack");
        }
        std::vector<uint8_t> dense(host->size());
        check(lgdhip_grid_store(grid->get(), &g, dense.data(), stream.get()), "lgdhip_grid_store");
        for (std::size_t i = 0; i < dense.size(); ++i) {
            host->data()[i].*MEMBER = static_cast<VALUE>(dense[i]);
        }
    }

    // byte-grid mode: the streak's bytes of the byte grid; otherwise the
    // packed words covering the streak, unpacked on the host
    void downloadRegion(Grid<CELL, Topology> *host, const Region<2>& region)
    {
        const Coord<2> dim = host->boundingBox().dimensions;
        const std::vector<Streak<2> > streaks = relativeStreaks(region, host->boundingBox());
        const int64_t nw = (g.nx + 31) / 32;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<2>& st = streaks[k];
            const int y = st.origin.y();
            CELL *row = host->data() + (std::size_t)y * dim.x();
            if (bytes) {
                std::vector<uint8_t> line(st.endX - st.origin.x());
                const uint8_t *src = (const uint8_t *)grid->get() + elementIndex(g, 0, st.origin.x(), y, 0);
                check(lgdhip_memcpy_d2h(line.data(), src, line.size(), stream.get()), "lgdhip_memcpy_d2h");
                for (std::size_t i = 0; i < line.size(); ++i) {
                    row[st.origin.x() + i].*MEMBER = static_cast<VALUE>(line[i]);
                }
                continue;
            }
            const int64_t w0 = st.origin.x() / 32, w1 = (st.endX - 1) / 32;
            std::vector<uint32_t> words(w1 - w0 + 1);
            check(lgdhip_memcpy_d2h(words.data(), pa + y * nw + w0, words.size() * 4, stream.get()),
                  "lgdhip_memcpy_d2h");
            for (int x = st.origin.x(); x < st.endX; ++x) {
                row[x].*MEMBER = static_cast<VALUE>((words[x / 32 - w0] >> (x % 32)) & 1u);
            }
        }
    }

    void run(unsigned nanoSteps)
    {
        if (bytes) {
            // halo refresh + one generation per step, two steps per
            // captured graph
            void *pa = grid->get(), *pb = grid2->get();
            loop.run(stream.get(), pa, nanoSteps / 2, [&]() {
                wrappedGeneration(pa, pb);
                wrappedGeneration(pb, pa);
            });
            if (nanoSteps % 2) {
                wrappedGeneration(pa, pb);
                grid.swap(grid2);
            }
            return;
        }
        int inB = 0;
        if (wrap) {
            check(lgdhip_life2d_packed_run_periodic(pa, pb, &g, edgeAlive, (int)nanoSteps, wrap, tuning.get(), &inB,
                                                    stream.get()),
                  "lgdhip_life2d_packed_run_periodic");
        } else {
            check(lgdhip_life2d_packed_run(pa, pb, &g, edgeAlive, (int)nanoSteps, tuning.get(), &inB, stream.get()),
                  "lgdhip_life2d_packed_run");
        }
        if (inB) {
            std::swap(pa, pb);
        }
    }

    void sync() { stream.sync(); }
    const GraphLoop& graphs() const { return loop; }

private:
    Tuning tuning;
    Stream stream;
    lgdhip_grid g;
    std::unique_ptr<DeviceMemory> grid, grid2, packed;
    uint32_t *pa, *pb;
    int edgeAlive;
    int wrap;
    bool bytes;  // byte-grid single generations (wrapped x with nx % 32 != 0)
    GraphLoop loop;

    void wrappedGeneration(void *src, void *dst)
    {
        const uint8_t edge = (uint8_t)edgeAlive;
        int inB = 0;
        check(lgdhip_grid_fill_halo(src, &g, &edge, wrap, stream.get()), "lgdhip_grid_fill_halo");
        check(lgdhip_life2d_u8_run((uint8_t *)src, (uint8_t *)dst, &g, 1, &inB, stream.get()),
              "lgdhip_life2d_u8_run");
    }
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

    void uploadRegion(const Grid<CELL, Topology>& host, const Region<3>& region)
    {
        const Coord<3> dim = host.boundingBox().dimensions;
        const std::vector<Streak<3> > streaks = relativeStreaks(region, host.boundingBox());
        std::vector<float> line;
        std::vector<uint8_t> fl;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<3>& st = streaks[k];
            const std::size_t n = st.endX - st.origin.x();
            const CELL *cells = host.data() + st.origin.toIndex(dim);
            line.resize(n);
            for (int q = 0; q < 19; ++q) {
                for (std::size_t i = 0; i < n; ++i) {
                    line[i] = (float)((double)cells[i].f[q] - weight(q));
                }
                float *dst = (float *)a->get() + elementIndex(g, q, st.origin.x(), st.origin.y(), st.origin.z());
                check(lgdhip_memcpy_h2d(dst, line.data(), n * sizeof(float), stream.get()), "lgdhip_memcpy_h2d");
            }
            fl.resize(n);
            for (std::size_t i = 0; i < n; ++i) {
                fl[i] = cells[i].flags;
            }
            uint8_t *fdst = (uint8_t *)flags->get() + elementIndex(fg, 0, st.origin.x(), st.origin.y(), st.origin.z());
            check(lgdhip_memcpy_h2d(fdst, fl.data(), n, stream.get()), "lgdhip_memcpy_h2d");
        }
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

    // one strided copy per streak: the streak's segment of all 19 members
    void downloadRegion(Grid<CELL, Topology> *host, const Region<3>& region)
    {
        const Coord<3> dim = host->boundingBox().dimensions;
        const std::vector<Streak<3> > streaks = relativeStreaks(region, host->boundingBox());
        const std::size_t memberStride = (std::size_t)(elementIndex(g, 1, 0, 0, 0) - elementIndex(g, 0, 0, 0, 0));
        std::vector<float> line;
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Streak<3>& st = streaks[k];
            const std::size_t n = st.endX - st.origin.x();
            line.resize(19 * n);
            const float *src = (const float *)a->get() + elementIndex(g, 0, st.origin.x(), st.origin.y(),
                                                                         st.origin.z());
            check(lgdhip_memcpy2d_d2h(line.data(), n * sizeof(float), src, memberStride * sizeof(float),
                                      n * sizeof(float), 19, stream.get()),
                  "lgdhip_memcpy2d_d2h");
            CELL *cells = host->data() + st.origin.toIndex(dim);
            for (std::size_t i = 0; i < n; ++i) {
                for (int q = 0; q < 19; ++q) {
                    cells[i].f[q] = (float)((double)line[q * n + i] + weight(q));
                }
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
