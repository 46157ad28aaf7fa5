// Initializer<CELL>: sets up the initial grid and the simulation's time
// range (src/libgeodecomp/io/initializer.h:38-71): grid(), gridBox(),
// gridDimensions(), startStep(), maxSteps().
#ifndef LIBGEODECOMP_IO_INITIALIZER_H
#define LIBGEODECOMP_IO_INITIALIZER_H

#include <libgeodecomp/misc/apitraits.h>
#include <libgeodecomp/storage/gridbase.h>

namespace LibGeoDecomp {

template<typename CELL>
class Initializer
{
public:
    typedef typename APITraits::SelectTopology<CELL>::Value Topology;
    typedef CELL Cell;
    static const unsigned NANO_STEPS = APITraits::SelectNanoSteps<CELL>::VALUE;
    static const int DIM = Topology::DIM;

    virtual ~Initializer() {}

    // initializes all cells of the grid at target; may be called repeatedly
    virtual void grid(GridBase<CELL, DIM> *target) = 0;

    virtual CoordBox<DIM> gridBox() { return CoordBox<DIM>(Coord<DIM>(), gridDimensions()); }
    virtual Coord<DIM> gridDimensions() const = 0;
    virtual unsigned startStep() const = 0;
    // the simulator updates from startStep() until the step counter equals maxSteps()
    virtual unsigned maxSteps() const = 0;

    Coord<DIM> normalize(const Coord<DIM>& coord) const
    {
        return Topology::normalize(coord, gridDimensions());
    }
};

}

#endif
