// MonolithicSimulator<CELL>: a Simulator whose whole grid lives in one
// place and which serves Writers (src/libgeodecomp/parallelization/
// monolithicsimulator.h:17-52).  It owns every Writer passed to it.
#ifndef LIBGEODECOMP_PARALLELIZATION_MONOLITHICSIMULATOR_H
#define LIBGEODECOMP_PARALLELIZATION_MONOLITHICSIMULATOR_H

#include <libgeodecomp/io/writer.h>
#include <libgeodecomp/parallelization/simulator.h>

namespace LibGeoDecomp {

template<typename CELL_TYPE>
class MonolithicSimulator : public Simulator<CELL_TYPE>
{
public:
    typedef typename Simulator<CELL_TYPE>::Topology Topology;
    typedef std::shared_ptr<Writer<CELL_TYPE> > WriterPtr;
    typedef std::vector<WriterPtr> WriterVector;
    using Simulator<CELL_TYPE>::chronometer;

    explicit MonolithicSimulator(Initializer<CELL_TYPE> *initializer) :
        Simulator<CELL_TYPE>(initializer)
    {}

    virtual const typename Simulator<CELL_TYPE>::GridType *getGrid() = 0;

    virtual void addWriter(Writer<CELL_TYPE> *writer) { writers.push_back(WriterPtr(writer)); }

    std::vector<Chronometer> gatherStatistics() { return std::vector<Chronometer>(1, chronometer); }

protected:
    WriterVector writers;
};

}

#endif
