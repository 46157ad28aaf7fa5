// HipSlabStepper<CELL>: the inner-set / rim split of the reference's
// CUDAStepper (src/libgeodecomp/parallelization/nesting/cudastepper.h:255-590,
// update1 :336-380, updateGhost :516-575) for cells of the Jacobi3DF64
// kernel family, on one GPU.  The grid is cut into z slabs; each slab has
// its own device buffers holding its planes plus a ghost zone of
// ghostZoneWidth() planes towards each neighbouring slab.  One round
// advances every slab by up to ghostZoneWidth() steps in ONE temporally
// blocked launch over the slab and its ghost zones (the planes outside the
// slab's box are taken as the edge, which corrupts only the outermost K
// planes of the ghost zones after K sweeps: exactly what the next exchange
// overwrites).  Within a round the launch is split by z range: the rim
// (the ghostZoneWidth() planes next to each neighbour) runs on a rim stream
// and its planes go to the neighbours' ghost zones on a copy stream, ordered
// after the rim by an event, while the inner planes are computed on the
// compute stream.  Rim and inner launches of one round read the same buffer
// and write disjoint planes, so they overlap; the next round's inner
// launches wait for this round's copies (which waited for the rims), its
// rims for this round's inner launches and copies.
// This is the step loop a multi-GPU run would use with the copies crossing
// to other GPUs; here all slabs share one GPU, so the copies are
// device-to-device.  Results equal HipSimulator's (and the oracle's) bit
// for bit.
//
// The reference's stepper takes its rim and inner sets from a
// PartitionManager and hands ghost zones to PatchAccepters/Providers of an
// MPI update group; here the partition is the slab list and the "patch
// link" the copy stream (one process, one GPU).  Cube topology only.
#ifndef LIBGEODECOMP_PARALLELIZATION_NESTING_HIPSLABSTEPPER_H
#define LIBGEODECOMP_PARALLELIZATION_NESTING_HIPSLABSTEPPER_H

#include <libgeodecomp/hip/kernelfamilies.h>
#include <libgeodecomp/misc/apitraits.h>
#include <libgeodecomp/storage/grid.h>

#include <lgdhip.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

namespace LibGeoDecomp {

namespace HipKernels {

// the FP64 member a Jacobi3DF64 family sweeps
template<typename FAMILY>
struct SlabMember;

template<typename CELL, typename VALUE, VALUE CELL::*MEMBER>
struct SlabMember<Jacobi3DF64<CELL, VALUE, MEMBER> >
{
    static double get(const CELL& c) { return c.*MEMBER; }
    static void set(CELL *c, double v) { c->*MEMBER = v; }
};

class Event
{
public:
    Event() : e(0) { check(lgdhip_event_create(&e), "lgdhip_event_create"); }
    ~Event() { lgdhip_event_destroy(e); }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    lgdhip_event_t get() const { return e; }

private:
    lgdhip_event_t e;
};

}

template<typename CELL>
class HipSlabStepper
{
public:
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    typedef typename APITraits::SelectHipKernel<CELL>::Value Family;
    typedef HipKernels::SlabMember<Family> Member;  // fails to compile for other families
    typedef Grid<CELL, Topology> GridType;
    static const int DIM = 3;

    // slabs: number of z slabs (about equal thickness, each at least
    // ghostZoneWidth planes); ghostZoneWidth: planes exchanged per neighbour
    // and round, and the most steps one round advances (1-4, a
    // temporal-blocking depth)
    HipSlabStepper(const GridType& initial, int slabs, int ghostZoneWidth = 4) :
        width(ghostZoneWidth),
        step(0),
        planesCopied(0),
        rimLaunches(0),
        innerLaunches(0),
        rounds(0)
    {
        for (int d = 0; d < DIM; ++d) {
            if (Topology::wrapsAxis(d)) {
                throw std::logic_error("HipSlabStepper implements the Cube topology only");
            }
        }
        const Coord<3> dim = initial.boundingBox().dimensions;
        if (width < 1 || width > 4) {
            throw std::invalid_argument("ghost zone width must be in [1, 4]");
        }
        // a ghost zone comes from the adjacent slab alone: with several
        // slabs each must be at least as thick as the ghost zone
        if (slabs < 1 || (slabs > 1 && dim.z() / slabs < width)) {
            throw std::invalid_argument("need slabs >= 1 and, with several slabs, nz / slabs >= ghost zone width");
        }
        nx = dim.x();
        ny = dim.y();
        edge = Member::get(initial.getEdge());
        for (int s = 0; s < slabs; ++s) {
            Slab slab;
            slab.z0 = (int)((int64_t)dim.z() * s / slabs);
            slab.z1 = (int)((int64_t)dim.z() * (s + 1) / slabs);
            slab.lo = s > 0 ? width : 0;
            slab.hi = s + 1 < slabs ? width : 0;
            const int planes = slab.z1 - slab.z0 + slab.lo + slab.hi;
            HipKernels::check(lgdhip_grid_make(&slab.g, 3, nx, ny, planes, 1, 8, 1), "lgdhip_grid_make");
            const std::size_t bytes = (std::size_t)lgdhip_grid_bytes(&slab.g);
            slab.cur.reset(new HipKernels::DeviceMemory(bytes));
            slab.nxt.reset(new HipKernels::DeviceMemory(bytes));
            slab.rimDone.reset(new HipKernels::Event());
            // own planes and ghost zones from the initial grid
            std::vector<double> dense((std::size_t)nx * ny * planes);
            for (int z = 0; z < planes; ++z) {
                for (int y = 0; y < ny; ++y) {
                    for (int x = 0; x < nx; ++x) {
                        dense[((std::size_t)z * ny + y) * nx + x] =
                            Member::get(initial.get(Coord<3>(x, y, slab.z0 - slab.lo + z)));
                    }
                }
            }
            for (void *p : {slab.cur->get(), slab.nxt->get()}) {
                HipKernels::check(lgdhip_grid_fill_edge(p, &slab.g, &edge, compute.get()), "lgdhip_grid_fill_edge");
            }
            HipKernels::check(lgdhip_grid_load(slab.cur->get(), &slab.g, dense.data(), compute.get()),
                              "lgdhip_grid_load");
            slabList.push_back(std::move(slab));
        }
        compute.sync();
    }

    ~HipSlabStepper()
    {
        // the buffers outlive the streams' queued work
        lgdhip_stream_synchronize(compute.get());
        lgdhip_stream_synchronize(rims.get());
        lgdhip_stream_synchronize(copies.get());
    }

    // advance every slab by `steps` steps: rounds of ghostZoneWidth() steps
    // (the last one shorter if steps is not a multiple)
    void update(unsigned steps)
    {
        while (steps > 0) {
            const int k = (int)std::min<unsigned>(steps, (unsigned)width);
            round(k);
            steps -= (unsigned)k;
            step += (unsigned)k;
        }
    }

    // the current state of every slab's own planes
    void grid(GridType *out)
    {
        sync();
        for (Slab& slab : slabList) {
            std::vector<double> dense((std::size_t)nx * ny * (slab.z1 - slab.z0 + slab.lo + slab.hi));
            HipKernels::check(lgdhip_grid_store(slab.cur->get(), &slab.g, dense.data(), compute.get()),
                              "lgdhip_grid_store");
            for (int z = slab.z0; z < slab.z1; ++z) {
                for (int y = 0; y < ny; ++y) {
                    for (int x = 0; x < nx; ++x) {
                        CELL c = out->get(Coord<3>(x, y, z));
                        Member::set(&c, dense[((std::size_t)(z - slab.z0 + slab.lo) * ny + y) * nx + x]);
                        out->set(Coord<3>(x, y, z), c);
                    }
                }
            }
        }
    }

    // wait for every queued round (for timing; grid() waits by itself)
    void sync()
    {
        compute.sync();
        rims.sync();
        copies.sync();
    }

    unsigned currentStep() const { return step; }
    int ghostZoneWidth() const { return width; }
    std::size_t slabs() const { return slabList.size(); }
    // the z range [first, second) slab s owns
    std::pair<int, int> slabRange(std::size_t s) const { return std::make_pair(slabList[s].z0, slabList[s].z1); }
    unsigned long long planesExchanged() const { return planesCopied; }
    unsigned long long rimUpdates() const { return rimLaunches; }
    unsigned long long innerUpdates() const { return innerLaunches; }
    unsigned long long roundsRun() const { return rounds; }

private:
    struct Slab
    {
        int z0, z1;  // owned global planes
        int lo, hi;  // ghost planes below and above
        lgdhip_grid g;
        std::unique_ptr<HipKernels::DeviceMemory> cur, nxt;
        std::unique_ptr<HipKernels::Event> rimDone;
    };

    int width;
    int nx, ny;
    double edge;
    unsigned step;
    unsigned long long planesCopied, rimLaunches, innerLaunches, rounds;
    std::vector<Slab> slabList;
    HipKernels::Stream compute, rims, copies;
    HipKernels::Event copiesDone, innerDone;

    // variant of the blocked kernels for k sweeps per launch (4: the
    // default plan's choice for the slab's shape)
    static int variantFor(int k) { return k == 1 ? 11 : (k == 2 ? 4 : (k == 3 ? 5 : 0)); }

    void launch(Slab& slab, int z0, int z1, int k, lgdhip_stream_t stream, unsigned long long *counter)
    {
        if (z0 >= z1) {
            return;
        }
        lgdhip_tile tile;
        std::memset(&tile, 0, sizeof(tile));
        tile.variant = variantFor(k);
        int sweeps = 0;
        HipKernels::check(lgdhip_jacobi3d_f64_run_zrange((const double *)slab.cur->get(), (double *)slab.nxt->get(),
                                                         &slab.g, z0, z1, &tile, &sweeps, stream),
                          "lgdhip_jacobi3d_f64_run_zrange");
        if (sweeps != k) {
            throw std::logic_error("blocked launch ran an unexpected number of sweeps");
        }
        ++*counter;
    }

    // `count` planes from slab-local plane `from` of src's next buffer to
    // plane `to` of dst's next buffer
    void copyPlanes(Slab& src, int from, Slab& dst, int to, int count)
    {
        const std::size_t plane = (std::size_t)src.g.sx * src.g.sy;
        const double *s = (const double *)src.nxt->get() + (std::size_t)(from + src.g.oz) * plane;
        double *d = (double *)dst.nxt->get() + (std::size_t)(to + dst.g.oz) * plane;
        HipKernels::check(lgdhip_memcpy_d2d(d, s, (std::size_t)count * plane * sizeof(double), copies.get()),
                          "lgdhip_memcpy_d2d");
        planesCopied += (unsigned long long)count;
    }

    // k <= width steps on every slab: rim planes (rim stream) and their
    // copies to the neighbours' ghost zones (copy stream), next to the inner
    // planes (compute stream)
    void round(int k)
    {
        const std::size_t n = slabList.size();
        // the rims read the ghost zones the last round's copies wrote and the
        // planes its inner launches wrote, and overwrite what those read
        if (rounds > 0) {
            HipKernels::check(lgdhip_stream_wait_event(rims.get(), copiesDone.get()), "lgdhip_stream_wait_event");
            HipKernels::check(lgdhip_stream_wait_event(rims.get(), innerDone.get()), "lgdhip_stream_wait_event");
        }
        for (std::size_t s = 0; s < n; ++s) {
            Slab& slab = slabList[s];
            const int own0 = slab.lo, own1 = slab.lo + (slab.z1 - slab.z0);
            // rim: the planes each neighbour keeps as its ghost zone
            const int rimLo = s > 0 ? std::min(own0 + slabList[s - 1].hi, own1) : own0;
            const int rimHi = s + 1 < n ? std::max(own1 - slabList[s + 1].lo, rimLo) : own1;
            launch(slab, own0, rimLo, k, rims.get(), &rimLaunches);
            if (rimHi < own1) {
                launch(slab, std::max(rimHi, rimLo), own1, k, rims.get(), &rimLaunches);
            }
            HipKernels::check(lgdhip_event_record(slab.rimDone->get(), rims.get()), "lgdhip_event_record");
            HipKernels::check(lgdhip_stream_wait_event(copies.get(), slab.rimDone->get()),
                              "lgdhip_stream_wait_event");
            if (s > 0) {
                Slab& below = slabList[s - 1];
                copyPlanes(slab, own0, below, below.lo + (below.z1 - below.z0), below.hi);
            }
            if (s + 1 < n) {
                Slab& above = slabList[s + 1];
                copyPlanes(slab, own1 - above.lo, above, 0, above.lo);
            }
            // inner planes while the rims and copies run
            launch(slab, rimLo, std::max(rimHi, rimLo), k, compute.get(), &innerLaunches);
        }
        // the next round's inner launches read the ghost zones and rim
        // planes (the copies waited for every rim) and overwrite what the
        // rims read
        HipKernels::check(lgdhip_event_record(copiesDone.get(), copies.get()), "lgdhip_event_record");
        HipKernels::check(lgdhip_event_record(innerDone.get(), compute.get()), "lgdhip_event_record");
        HipKernels::check(lgdhip_stream_wait_event(compute.get(), copiesDone.get()), "lgdhip_stream_wait_event");
        for (Slab& slab : slabList) {
            std::swap(slab.cur, slab.nxt);
        }
        ++rounds;
    }
};

}

#endif
