// SimpleInitializer<CELL>: fixed dimensions and step count, start step 0
// (src/libgeodecomp/io/simpleinitializer.h).
#ifndef LIBGEODECOMP_IO_SIMPLEINITIALIZER_H
#define LIBGEODECOMP_IO_SIMPLEINITIALIZER_H

#include <libgeodecomp/io/initializer.h>

namespace LibGeoDecomp {

template<typename CELL>
class SimpleInitializer : public Initializer<CELL>
{
public:
    typedef typename Initializer<CELL>::Topology Topology;
    const static int DIM = Topology::DIM;

    explicit SimpleInitializer(const Coord<DIM>& dimensions, const unsigned steps = 300) :
        dimensions(dimensions),
        steps(steps)
    {}

    Coord<DIM> gridDimensions() const { return dimensions; }
    unsigned maxSteps() const { return steps; }
    unsigned startStep() const { return 0; }

protected:
    Coord<DIM> dimensions;
    unsigned steps;
};

}

#endif
