// Simulator<CELL>: abstract step loop with Initializer and Steerers
// (src/libgeodecomp/parallelization/simulator.h:28-105).  The simulator
// owns the Initializer and every Steerer passed to it.
#ifndef LIBGEODECOMP_PARALLELIZATION_SIMULATOR_H
#define LIBGEODECOMP_PARALLELIZATION_SIMULATOR_H

#include <libgeodecomp/io/initializer.h>
#include <libgeodecomp/io/steerer.h>
#include <libgeodecomp/misc/chronometer.h>

#include <memory>
#include <vector>

namespace LibGeoDecomp {

template<typename CELL_TYPE>
class Simulator
{
public:
    typedef typename APITraits::SelectTopology<CELL_TYPE>::Value Topology;
    static const int DIM = Topology::DIM;
    static const unsigned NANO_STEPS = APITraits::SelectNanoSteps<CELL_TYPE>::VALUE;
    typedef GridBase<CELL_TYPE, DIM> GridType;
    typedef std::shared_ptr<Initializer<CELL_TYPE> > InitPtr;
    typedef std::shared_ptr<Steerer<CELL_TYPE> > SteererPtr;
    typedef std::vector<SteererPtr> SteererVector;

    explicit Simulator(Initializer<CELL_TYPE> *initializer) :
        stepNum(0),
        initializer(initializer),
        gridDim(initializer->gridDimensions())
    {}

    virtual ~Simulator() {}

    virtual void step() = 0;
    virtual void run() = 0;
    virtual unsigned getStep() const { return stepNum; }
    virtual InitPtr getInitializer() const { return initializer; }

    virtual void addSteerer(Steerer<CELL_TYPE> *steerer) { steerers.push_back(SteererPtr(steerer)); }

    virtual std::vector<Chronometer> gatherStatistics() = 0;

protected:
    Chronometer chronometer;
    unsigned stepNum;
    InitPtr initializer;
    SteererVector steerers;
    Coord<DIM> gridDim;
};

}

#endif
