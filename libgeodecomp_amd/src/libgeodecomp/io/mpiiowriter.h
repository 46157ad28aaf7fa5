// MPIIOWriter<CELL>: periodic grid snapshots in the reference's MPI-IO
// format (src/libgeodecomp/io/mpiiowriter.h:18-98): the whole grid at the
// initial step, every `period` steps and the last step, to
// PREFIX%05d.mpiio (no separator before the step number), with maxSteps in
// the header so that MPIIOInitializer can resume the run.  One process,
// plain file I/O (see libgeodecomp/io/mpiio.h); the constructor keeps the
// reference's (prefix, period, maxSteps) arguments.
#ifndef LIBGEODECOMP_IO_MPIIOWRITER_H
#define LIBGEODECOMP_IO_MPIIOWRITER_H

#include <libgeodecomp/io/mpiio.h>
#include <libgeodecomp/io/writer.h>

#include <iomanip>
#include <sstream>

namespace LibGeoDecomp {

template<typename CELL_TYPE>
class MPIIOWriter : public Writer<CELL_TYPE>
{
public:
    typedef typename Writer<CELL_TYPE>::GridType GridType;
    typedef typename Writer<CELL_TYPE>::Topology Topology;
    static const int DIM = Topology::DIM;
    using Writer<CELL_TYPE>::period;
    using Writer<CELL_TYPE>::prefix;

    MPIIOWriter(const std::string& prefix, const unsigned period, const unsigned maxSteps) :
        Writer<CELL_TYPE>(prefix, period),
        maxSteps(maxSteps)
    {}

    Writer<CELL_TYPE> *clone() const { return new MPIIOWriter(*this); }

    void stepFinished(const GridType& grid, unsigned step, WriterEvent event)
    {
        if (event == WRITER_STEP_FINISHED && step % period != 0) {
            return;
        }
        Region<DIM> region;
        region << grid.boundingBox();
        mpiio.writeRegion(grid, grid.dimensions(), step, maxSteps, filename(step), region);
    }

    std::string filename(unsigned step) const
    {
        std::ostringstream buf;
        buf << prefix << std::setfill('0') << std::setw(5) << step << ".mpiio";
        return buf.str();
    }

private:
    unsigned maxSteps;
    MPIIO<CELL_TYPE, Topology> mpiio;
};

}

#endif
