// MPIIOInitializer<CELL>: resumes a run from a snapshot in the reference's
// MPI-IO format (src/libgeodecomp/io/mpiioinitializer.h:16-78): grid
// dimensions, startStep() and maxSteps() come from the file's header, the
// cells and the edge cell from its body.  Simulators start counting at
// startStep(), so writers and steerers keep their phase across a restart.
#ifndef LIBGEODECOMP_IO_MPIIOINITIALIZER_H
#define LIBGEODECOMP_IO_MPIIOINITIALIZER_H

#include <libgeodecomp/io/initializer.h>
#include <libgeodecomp/io/mpiio.h>

namespace LibGeoDecomp {

template<typename CELL_TYPE>
class MPIIOInitializer : public Initializer<CELL_TYPE>
{
public:
    typedef typename APITraits::SelectTopology<CELL_TYPE>::Value Topology;
    static const int DIM = Topology::DIM;

    explicit MPIIOInitializer(const std::string& filename) :
        file(filename)
    {
        mpiio.readMetadata(&dimensions, &currentStep, &maximumSteps, file);
    }

    virtual void grid(GridBase<CELL_TYPE, DIM> *target)
    {
        Region<DIM> region;
        region << target->boundingBox();
        mpiio.readRegion(target, file, region);
    }

    virtual Coord<DIM> gridDimensions() const { return dimensions; }
    virtual unsigned maxSteps() const { return maximumSteps; }
    virtual unsigned startStep() const { return currentStep; }

private:
    std::string file;
    MPIIO<CELL_TYPE, Topology> mpiio;
    unsigned currentStep;
    unsigned maximumSteps;
    Coord<DIM> dimensions;
};

}

#endif
