// User cells of the struct-of-arrays path: registered with
// LIBFLATARRAY_REGISTER_SOA and updated by a static
// updateLineX(hoodOld, indexEnd, hoodNew, nanoStep) -- the cell concept of
// the reference's FixedNeighborhoodUpdateFunctor
// (storage/fixedneighborhoodupdatefunctor.h:212-253) -- plus an
// array-of-structs line cell (updateLineX(target, &x, endX, hood, nanoStep),
// storage/linepointerupdatefunctor.h:162-174).  The same cells run through
// the oracle SerialSimulator (g++) and HipSimulator (hipcc).
#ifndef LGD_TEST_SOACELLS_H
#define LGD_TEST_SOACELLS_H

#include "devicecells.h"

#include <libflatarray/flat_array.hpp>
#include <libgeodecomp/misc/hostdevice.h>

// ---------------------------------------------------------------------
// Jacobi 7-point in SoA form: the reference's SoA streak cell
// (src/testbed/performancetests/main.cpp:1274-1435) with a scalar loop in
// place of its SSE one and the 7-point sum in JacobiCellFixedHood's order
// (main.cpp:1196-1205), so it matches the hand-written family bit for bit.
// ---------------------------------------------------------------------
template<typename TOPOLOGY>
class SoAJacobiCellOn
{
public:
    class API :
        public APITraits::HasFixedCoordsOnlyUpdate,
        public APITraits::HasUpdateLineX,
        public APITraits::HasStencil<Stencils::VonNeumann<3, 1> >,
        public APITraits::HasTopology<TOPOLOGY>,
        public APITraits::HasSoA
    {};

    __host__ __device__ explicit SoAJacobiCellOn(double t = 0) : temp(t) {}

    template<typename HOOD_OLD, typename HOOD_NEW>
    __host__ __device__ static void updateLineX(HOOD_OLD& hoodOld, int indexEnd, HOOD_NEW& hoodNew, int /* nano */)
    {
        for (; hoodOld.index() < indexEnd; ++hoodOld.index(), ++hoodNew.index()) {
            hoodNew.temp() = (hoodOld[FixedCoord< 0,  0, -1>()].temp() +
                              hoodOld[FixedCoord< 0, -1,  0>()].temp() +
                              hoodOld[FixedCoord<-1,  0,  0>()].temp() +
                              hoodOld[FixedCoord< 0,  0,  0>()].temp() +
                              hoodOld[FixedCoord< 1,  0,  0>()].temp() +
                              hoodOld[FixedCoord< 0,  1,  0>()].temp() +
                              hoodOld[FixedCoord< 0,  0,  1>()].temp()) * (1.0 / 7.0);
        }
    }

    bool operator==(const SoAJacobiCellOn& o) const { return temp == o.temp; }

    double temp;
};

typedef SoAJacobiCellOn<Topologies::Cube<3>::Topology> SoAJacobiCell;
typedef SoAJacobiCellOn<Topologies::Torus<3>::Topology> SoATorusJacobiCell;
typedef SoAJacobiCellOn<TopologiesHelpers::Topology<3, true, false, true> > SoASlabJacobiCell;

LIBFLATARRAY_REGISTER_SOA(SoAJacobiCell, ((double)(temp)))
LIBFLATARRAY_REGISTER_SOA(SoATorusJacobiCell, ((double)(temp)))
LIBFLATARRAY_REGISTER_SOA(SoASlabJacobiCell, ((double)(temp)))

// ---------------------------------------------------------------------
// Two-member cells of the reference's functor test
// (storage/test/unit/fixedneighborhoodupdatefunctortest.h:10-100): valB =
// the sum of the six face neighbours' valA, valA left alone.
// ---------------------------------------------------------------------
template<typename TOPOLOGY>
class TwoDoublesCell
{
public:
    class API :
        public APITraits::HasFixedCoordsOnlyUpdate,
        public APITraits::HasUpdateLineX,
        public APITraits::HasStencil<Stencils::Moore<3, 1> >,
        public APITraits::HasTopology<TOPOLOGY>,
        public APITraits::HasSoA
    {};

    __host__ __device__ explicit TwoDoublesCell(double valA = 0, double valB = 0) : valA(valA), valB(valB) {}

    template<typename HOOD_OLD, typename HOOD_NEW>
    __host__ __device__ static void updateLineX(HOOD_OLD& hoodOld, long indexEnd, HOOD_NEW& hoodNew, long /* nano */)
    {
        for (; hoodOld.index() < indexEnd; hoodOld += 1, hoodNew += 1) {
            double akku = hoodOld[FixedCoord< 0,  0, -1>()].valA() +
                          hoodOld[FixedCoord< 0, -1,  0>()].valA() +
                          hoodOld[FixedCoord<-1,  0,  0>()].valA() +
                          hoodOld[FixedCoord< 1,  0,  0>()].valA() +
                          hoodOld[FixedCoord< 0,  1,  0>()].valA() +
                          hoodOld[FixedCoord< 0,  0,  1>()].valA();
            hoodNew.valB() = akku;
        }
    }

    double valA;
    double valB;
};

typedef TwoDoublesCell<Topologies::Torus<3>::Topology> TwoDoublesTorus;
typedef TwoDoublesCell<Topologies::Cube<3>::Topology> TwoDoublesCube;

LIBFLATARRAY_REGISTER_SOA(TwoDoublesTorus, ((double)(valA))((double)(valB)))
LIBFLATARRAY_REGISTER_SOA(TwoDoublesCube, ((double)(valA))((double)(valB)))

// every cell valA = x + 1000 y + 10^6 z (the functor test's encoding), the
// edge cell valA = edgeA, valB = -1 everywhere
template<typename CELL>
class EncodedInitializer : public SimpleInitializer<CELL>
{
public:
    EncodedInitializer(const Coord<3>& dim, unsigned steps, double edgeA) :
        SimpleInitializer<CELL>(dim, steps), edgeA(edgeA)
    {}

    virtual void grid(GridBase<CELL, 3> *ret)
    {
        ret->setEdge(CELL(edgeA, -1));
        CoordBox<3> box = ret->boundingBox();
        for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
            ret->set(*i, CELL(i->x() + i->y() * 1000.0 + i->z() * 1000.0 * 1000.0, -1));
        }
    }

private:
    double edgeA;
};

// ---------------------------------------------------------------------
// SoA probe cell after the reference's TestCellSoA (misc/testcell.h:
// TestCellHelpers::SoAAPI, :272-281 updateLineX, :424-431 registration):
// every nano step it checks its whole Moore neighbourhood through
// FixedCoord accesses only -- the edge cell exactly where the topology puts
// it, otherwise a valid cell at the topology-normalised position with the
// same cycle count -- and the nano step against its cycle count; failures
// clear `valid` and set a bit of ProbeFailure in `bad`.  `value` averages
// the neighbourhood in a fixed order (an FP32 result the GPU reproduces bit
// for bit).  updateLineX takes both accessors by value, as TestCell does.
// ---------------------------------------------------------------------
template<int DIM, int K, int END>
class SoAMooreVisit
{
public:
    static const int X = K % 3 - 1;
    static const int Y = DIM > 1 ? (K / 3) % 3 - 1 : 0;
    static const int Z = DIM > 2 ? K / 9 - 1 : 0;

    template<typename CELL, typename HOOD>
    __host__ __device__ static void run(CELL *cell, const HOOD& hood, int *ok, float *sum)
    {
        cell->template check<X, Y, Z>(hood[FixedCoord<X, Y, Z>()], ok, sum);
        SoAMooreVisit<DIM, K + 1, END>::run(cell, hood, ok, sum);
    }
};

template<int DIM, int END>
class SoAMooreVisit<DIM, END, END>
{
public:
    template<typename CELL, typename HOOD>
    __host__ __device__ static void run(CELL *, const HOOD&, int *, float *)
    {}
};

template<int DIM, typename TOPOLOGY, unsigned NANO>
class SoAProbeCell
{
public:
    class API :
        public APITraits::HasTopology<TOPOLOGY>,
        public APITraits::HasNanoSteps<NANO>,
        public APITraits::HasStencil<Stencils::Moore<DIM, 1> >,
        public APITraits::HasSoA,
        public APITraits::HasFixedCoordsOnlyUpdate,
        public APITraits::HasUpdateLineX,
        public LibFlatArray::api_traits::has_default_3d_sizes_uniform
    {};

    static const int VOLUME = StencilsHelpers::Power<3, DIM>::VALUE;

    Coord<DIM> pos;
    Coord<DIM> dim;
    unsigned cycle;
    int edge;
    int valid;
    int bad;
    float value;

    __host__ __device__ SoAProbeCell() : cycle(0), edge(0), valid(0), bad(0), value(0) {}

    __host__ __device__ SoAProbeCell(const Coord<DIM>& pos, const Coord<DIM>& dim, float value) :
        pos(pos), dim(dim), cycle(0), edge(0), valid(1), bad(0), value(value)
    {}

    static SoAProbeCell makeEdge(const Coord<DIM>& dim)
    {
        SoAProbeCell c(Coord<DIM>::diagonal(-1), dim, -1.0f);
        c.edge = 1;
        return c;
    }

    template<typename HOOD>
    __host__ __device__ void update(const HOOD& hood, unsigned nanoStep)
    {
        *this = SoAProbeCell(hood[FixedCoord<0, 0, 0>()]);
        if (edge) {
            valid = 0;
            bad |= PROBE_UPDATED_EDGE;
            return;
        }
        int ok = valid;
        float sum = 0.0f;
        SoAMooreVisit<DIM, 0, VOLUME>::run(this, hood, &ok, &sum);
        if (nanoStep != cycle % NANO) {
            ok = 0;
            bad |= PROBE_NANO_STEP;
        }
        valid = ok;
        ++cycle;
        value = sum * (1.0f / (float)VOLUME);
    }

    template<int X, int Y, int Z>
    __host__ __device__ void check(const SoAProbeCell& n, int *ok, float *sum)
    {
        const Coord<DIM> raw = pos + Coord<DIM>(FixedCoord<X, Y, Z>());
        const int shouldBeEdge = TOPOLOGY::isOutOfBounds(raw, dim) ? 1 : 0;
        if (n.edge != shouldBeEdge) {
            *ok = 0;
            bad |= PROBE_BAD_EDGE;
        } else if (!shouldBeEdge) {
            if (!n.valid) {
                *ok = 0;
                bad |= PROBE_INVALID_NEIGHBOR;
            }
            if (n.cycle != cycle) {
                *ok = 0;
                bad |= PROBE_CYCLE;
            }
            if (n.pos != TOPOLOGY::normalize(raw, dim)) {
                *ok = 0;
                bad |= PROBE_POSITION;
            }
        }
        *sum += n.value;
    }

    template<typename ACCESSOR1, typename ACCESSOR2>
    __host__ __device__ static void updateLineX(ACCESSOR1 hoodOld, int indexEnd, ACCESSOR2 hoodNew,
                                                unsigned nanoStep)
    {
        for (; hoodOld.index() < indexEnd; ++hoodOld.index(), ++hoodNew.index()) {
            SoAProbeCell cell;
            cell.update(hoodOld, nanoStep);
            hoodNew << cell;
        }
    }

    bool operator==(const SoAProbeCell& o) const
    {
        return pos == o.pos && dim == o.dim && cycle == o.cycle && edge == o.edge && valid == o.valid &&
               bad == o.bad && value == o.value;
    }
};

typedef SoAProbeCell<3, Topologies::Cube<3>::Topology, 1> SoAProbe3Cube;
typedef SoAProbeCell<3, Topologies::Torus<3>::Topology, 3> SoAProbe3Torus;
typedef SoAProbeCell<3, TopologiesHelpers::Topology<3, false, true, true>, 2> SoAProbe3Tube;
typedef SoAProbeCell<2, Topologies::Cube<2>::Topology, 1> SoAProbe2Cube;
typedef SoAProbeCell<2, Topologies::Torus<2>::Topology, 2> SoAProbe2Torus;
typedef SoAProbeCell<1, Topologies::Torus<1>::Topology, 1> SoAProbe1Ring;

#define LGD_REGISTER_SOA_PROBE(CELL, D)                                                                       \
    LIBFLATARRAY_REGISTER_SOA(CELL, ((LibGeoDecomp::Coord<D>)(pos))((LibGeoDecomp::Coord<D>)(dim))            \
                                        ((unsigned)(cycle))((int)(edge))((int)(valid))((int)(bad))((float)(value)))
LGD_REGISTER_SOA_PROBE(SoAProbe3Cube, 3)
LGD_REGISTER_SOA_PROBE(SoAProbe3Torus, 3)
LGD_REGISTER_SOA_PROBE(SoAProbe3Tube, 3)
LGD_REGISTER_SOA_PROBE(SoAProbe2Cube, 2)
LGD_REGISTER_SOA_PROBE(SoAProbe2Torus, 2)
LGD_REGISTER_SOA_PROBE(SoAProbe1Ring, 1)

// ---------------------------------------------------------------------
// SoA cell with private members, an array member and mixed member sizes:
// LIBFLATARRAY_ACCESS, ((float)(f)(3)) and a bool before a double
// ---------------------------------------------------------------------
class SoAMixedCell
{
public:
    LIBFLATARRAY_ACCESS

    class API :
        public APITraits::HasFixedCoordsOnlyUpdate,
        public APITraits::HasUpdateLineX,
        public APITraits::HasStencil<Stencils::VonNeumann<2, 1> >,
        public APITraits::HasTorusTopology<2>,
        public APITraits::HasSoA
    {};

    __host__ __device__ explicit SoAMixedCell(float a = 0, bool on = false, double d = 0) : on(on), d(d)
    {
        f[0] = a;
        f[1] = 2 * a;
        f[2] = -a;
    }

    // f[k] <- f[k] + sum of the four face neighbours' f[(k + 1) % 3], on
    // toggles, d accumulates f[0] of the west neighbour
    template<typename HOOD_OLD, typename HOOD_NEW>
    __host__ __device__ static void updateLineX(HOOD_OLD& hoodOld, int indexEnd, HOOD_NEW& hoodNew, unsigned)
    {
        for (; hoodOld.index() < indexEnd; ++hoodOld.index(), ++hoodNew.index()) {
            for (int k = 0; k < 3; ++k) {
                const int j = (k + 1) % 3;
                hoodNew.f()[k] = hoodOld[FixedCoord<0, 0, 0>()].f()[k] +
                                 (hoodOld[FixedCoord<-1, 0, 0>()].f()[j] + hoodOld[FixedCoord<1, 0, 0>()].f()[j] +
                                  hoodOld[FixedCoord<0, -1, 0>()].f()[j] + hoodOld[FixedCoord<0, 1, 0>()].f()[j]) *
                                     0.125f;
            }
            hoodNew.on() = !hoodOld[FixedCoord<0, 0, 0>()].on();
            hoodNew.d() = hoodOld[FixedCoord<0, 0, 0>()].d() + hoodOld[FixedCoord<-1, 0, 0>()].template f<0>();
        }
    }

    bool operator==(const SoAMixedCell& o) const
    {
        return f[0] == o.f[0] && f[1] == o.f[1] && f[2] == o.f[2] && on == o.on && d == o.d;
    }

    float get(int k) const { return f[k]; }
    bool isOn() const { return on; }
    double total() const { return d; }

private:
    float f[3];
    bool on;
    double d;
};

LIBFLATARRAY_REGISTER_SOA(SoAMixedCell, ((float)(f)(3))((bool)(on))((double)(d)))

// ---------------------------------------------------------------------
// Array-of-structs line cell: the jacobi3dupdateline example's form
// (src/examples/jacobi3dupdateline/main.cpp:32-42), 7-point sum in the
// Jacobi cell's order; runs through the generic family's line path.
// ---------------------------------------------------------------------
class LineJacobiCell
{
public:
    class API :
        public APITraits::HasFixedCoordsOnlyUpdate,
        public APITraits::HasUpdateLineX,
        public APITraits::HasStencil<Stencils::VonNeumann<3, 1> >,
        public APITraits::HasCubeTopology<3>
    {};

    __host__ __device__ explicit LineJacobiCell(double v = 0) : temp(v) {}

    template<typename NEIGHBORHOOD>
    __host__ __device__ static void updateLineX(LineJacobiCell *target, long *x, long endX,
                                                const NEIGHBORHOOD& hood, int /* nanoStep */)
    {
        for (; *x < endX; ++*x) {
            target[*x].temp = (hood[FixedCoord< 0,  0, -1>()].temp +
                               hood[FixedCoord< 0, -1,  0>()].temp +
                               hood[FixedCoord<-1,  0,  0>()].temp +
                               hood[FixedCoord< 0,  0,  0>()].temp +
                               hood[FixedCoord< 1,  0,  0>()].temp +
                               hood[FixedCoord< 0,  1,  0>()].temp +
                               hood[FixedCoord< 0,  0,  1>()].temp) * (1.0 / 7.0);
        }
    }

    bool operator==(const LineJacobiCell& o) const { return temp == o.temp; }

    double temp;
};

#endif
