// User cells whose update() runs on the GPU through the generic kernel
// family (no hand-written family registered): written as a LibGeoDecomp
// user writes a CUDA cell, update() marked __host__ __device__.  The same
// cells run through the oracle SerialSimulator on the CPU.
#ifndef LGD_TEST_DEVICECELLS_H
#define LGD_TEST_DEVICECELLS_H

#include "cells.h"

#include <libgeodecomp/misc/hostdevice.h>

// ---------------------------------------------------------------------
// ProbeCell: checks its whole neighbourhood every nano step, after the
// reference's TestCell (src/libgeodecomp/misc/testcell.h:130-395): every
// neighbour in the stencil must be the edge cell exactly where the
// topology says so, and otherwise a valid cell at the topology-normalised
// position with the same cycle count; the nano step must follow the cycle
// count.  Failures clear `valid` and set a bit in `bad` (which check
// failed), so the host can assert on every cell after a run.  `value`
// averages the neighbourhood (edge included) in a fixed order, an FP32
// result the GPU must reproduce bit for bit.  CHECKED is the topology the
// checks assume (the cell's own by default; a different one plants a
// mistake for the CPU tests that show the checks catch it).
// ---------------------------------------------------------------------
enum ProbeFailure {
    PROBE_UPDATED_EDGE = 1, PROBE_BAD_EDGE = 2, PROBE_INVALID_NEIGHBOR = 4, PROBE_CYCLE = 8, PROBE_POSITION = 16,
    PROBE_NANO_STEP = 32
};

template<int DIM, typename TOPOLOGY, unsigned NANO, typename STENCIL, typename CHECKED = TOPOLOGY>
class ProbeCell
{
public:
    class API :
        public APITraits::HasTopology<TOPOLOGY>,
        public APITraits::HasNanoSteps<NANO>,
        public APITraits::HasStencil<STENCIL>
    {};

    static const int R = STENCIL::RADIUS;
    static const int SIDE = 2 * R + 1;
    static const int BOX = StencilsHelpers::Power<SIDE, DIM>::VALUE;
    // VonNeumann stencils have fewer points than the (2R+1)^DIM box
    static const bool MANHATTAN = STENCIL::VOLUME != BOX;

    Coord<DIM> pos;
    Coord<DIM> dim;
    unsigned cycle;
    int edge;
    int valid;
    int bad;
    float value;

    __host__ __device__ ProbeCell() : cycle(0), edge(0), valid(0), bad(0), value(0) {}

    __host__ __device__ ProbeCell(const Coord<DIM>& pos, const Coord<DIM>& dim, float value) :
        pos(pos), dim(dim), cycle(0), edge(0), valid(1), bad(0), value(value)
    {}

    static ProbeCell makeEdge(const Coord<DIM>& dim)
    {
        ProbeCell c(Coord<DIM>::diagonal(-1), dim, -1.0f);
        c.edge = 1;
        return c;
    }

    template<typename HOOD>
    __host__ __device__ void update(const HOOD& hood, unsigned nanoStep)
    {
        *this = hood[FixedCoord<0, 0, 0>()];
        if (edge) {
            valid = 0;
            bad |= PROBE_UPDATED_EDGE;
            return;
        }
        int ok = valid;
        float sum = 0.0f;
        for (int k = 0; k < BOX; ++k) {
            Coord<DIM> rel;
            int rest = k;
            int manhattan = 0;
            for (int d = 0; d < DIM; ++d) {
                rel[d] = rest % SIDE - R;
                rest /= SIDE;
                manhattan += rel[d] < 0 ? -rel[d] : rel[d];
            }
            if (MANHATTAN && manhattan > R) {
                continue;
            }
            const ProbeCell& n = hood[rel];
            const Coord<DIM> raw = pos + rel;
            const int shouldBeEdge = CHECKED::isOutOfBounds(raw, dim) ? 1 : 0;
            if (n.edge != shouldBeEdge) {
                ok = 0;
                bad |= PROBE_BAD_EDGE;
            } else if (!shouldBeEdge) {
                if (!n.valid) {
                    ok = 0;
                    bad |= PROBE_INVALID_NEIGHBOR;
                }
                if (n.cycle != cycle) {
                    ok = 0;
                    bad |= PROBE_CYCLE;
                }
                if (n.pos != CHECKED::normalize(raw, dim)) {
                    ok = 0;
                    bad |= PROBE_POSITION;
                }
            }
            sum += n.value;
        }
        if (nanoStep != cycle % NANO) {
            ok = 0;
            bad |= PROBE_NANO_STEP;
        }
        valid = ok;
        ++cycle;
        value = sum * (1.0f / (float)STENCIL::VOLUME);
    }

    bool operator==(const ProbeCell& o) const
    {
        return pos == o.pos && dim == o.dim && cycle == o.cycle && edge == o.edge && valid == o.valid &&
               bad == o.bad && value == o.value;
    }
};

template<typename CELL>
class ProbeInitializer : public SimpleInitializer<CELL>
{
public:
    static const int DIM = SimpleInitializer<CELL>::DIM;

    ProbeInitializer(const Coord<DIM>& dim, unsigned steps, uint64_t seed) :
        SimpleInitializer<CELL>(dim, steps), seed(seed)
    {}

    virtual void grid(GridBase<CELL, DIM> *ret)
    {
        ret->setEdge(CELL::makeEdge(this->dimensions));
        CoordBox<DIM> box = ret->boundingBox();
        for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
            ret->set(*i, CELL(*i, this->dimensions, (float)uniform01(seed, i->toIndex(this->dimensions))));
        }
    }

private:
    uint64_t seed;
};

// ---------------------------------------------------------------------
// The Life and Jacobi cells again, without a registered family: their
// update() runs through the generic kernel, so the two GPU paths can be
// compared with each other as well as with the oracle.
// ---------------------------------------------------------------------
class DeviceConwayCell
{
public:
    class API
    {};

    __host__ __device__ explicit DeviceConwayCell(bool alive = false) : alive(alive) {}

    template<typename COORD_MAP>
    __host__ __device__ void update(const COORD_MAP& hood, unsigned)
    {
        int living = 0;
        for (int y = -1; y < 2; ++y) {
            for (int x = -1; x < 2; ++x) {
                living += hood[Coord<2>(x, y)].alive;
            }
        }
        living -= hood[Coord<2>(0, 0)].alive;
        alive = hood[Coord<2>(0, 0)].alive;
        alive = alive ? (2 <= living && living <= 3) : (living == 3);
    }

    bool operator==(const DeviceConwayCell& o) const { return alive == o.alive; }

    bool alive;
};

class DeviceJacobiCell
{
public:
    class API :
        public APITraits::HasStencil<Stencils::VonNeumann<3, 1> >,
        public APITraits::HasCubeTopology<3>
    {};

    __host__ __device__ explicit DeviceJacobiCell(double t = 0) : temp(t) {}

    template<typename NEIGHBORHOOD>
    __host__ __device__ void update(const NEIGHBORHOOD& hood, int)
    {
        temp = (hood[FixedCoord< 0,  0, -1>()].temp +
                hood[FixedCoord< 0, -1,  0>()].temp +
                hood[FixedCoord<-1,  0,  0>()].temp +
                hood[FixedCoord< 0,  0,  0>()].temp +
                hood[FixedCoord< 1,  0,  0>()].temp +
                hood[FixedCoord< 0,  1,  0>()].temp +
                hood[FixedCoord< 0,  0,  1>()].temp) * (1.0 / 7.0);
    }

    bool operator==(const DeviceJacobiCell& o) const { return temp == o.temp; }

    double temp;
};

// ---------------------------------------------------------------------
// Separate device updates: update() on the host, updateHIP() / updateCUDA()
// on the GPU (storage/cudaupdatefunctor.h:13-42); `path` records which ran.
// ---------------------------------------------------------------------
template<typename SEPARATE>
class SeparateUpdateCell
{
public:
    class API : public SEPARATE, public APITraits::HasTorusTopology<2>
    {};

    __host__ __device__ explicit SeparateUpdateCell(int v = 0) : v(v), path(0) {}

    template<typename HOOD>
    void update(const HOOD& hood, unsigned)
    {
        v = step(hood);
        path = 1;
    }

    template<typename HOOD>
    __host__ __device__ void updateHIP(const HOOD& hood, unsigned)
    {
        v = step(hood);
        path = 2;
    }

    template<typename HOOD>
    __host__ __device__ void updateCUDA(const HOOD& hood, unsigned)
    {
        v = step(hood);
        path = 3;
    }

    bool operator==(const SeparateUpdateCell& o) const { return v == o.v && path == o.path; }

    int v;
    int path;

private:
    template<typename HOOD>
    __host__ __device__ static int step(const HOOD& hood)
    {
        return (hood[FixedCoord<-1, 0, 0>()].v * 3 + hood[FixedCoord<1, 0, 0>()].v * 5 +
                hood[FixedCoord<0, -1, 0>()].v * 7 + hood[FixedCoord<0, 1, 0>()].v * 11 +
                hood[FixedCoord<0, 0, 0>()].v) % 1000003;
    }
};

typedef SeparateUpdateCell<APITraits::HasSeparateHIPUpdate> HipUpdateCell;
typedef SeparateUpdateCell<APITraits::HasSeparateCUDAUpdate> CudaUpdateCell;

#endif
