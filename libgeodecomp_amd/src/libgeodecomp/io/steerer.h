// Steerer<CELL>: input plugin that may read and modify the grid before a
// step (src/libgeodecomp/io/steerer.h:19-135): STEERER_INITIALIZED,
// STEERER_NEXT_STEP (when step % period == 0), STEERER_ALL_DONE;
// SteererFeedback::endSimulation() stops the run.
#ifndef LIBGEODECOMP_IO_STEERER_H
#define LIBGEODECOMP_IO_STEERER_H

#include <libgeodecomp/geometry/region.h>
#include <libgeodecomp/misc/apitraits.h>
#include <libgeodecomp/storage/gridbase.h>

#include <cstddef>
#include <stdexcept>

namespace LibGeoDecomp {

enum SteererEvent {
    STEERER_INITIALIZED,
    STEERER_NEXT_STEP,
    STEERER_ALL_DONE
};

template<typename CELL_TYPE>
class Steerer
{
public:
    typedef typename APITraits::SelectStaticData<CELL_TYPE>::Value StaticData;
    typedef typename APITraits::SelectTopology<CELL_TYPE>::Value Topology;
    typedef GridBase<CELL_TYPE, Topology::DIM> GridType;
    typedef Coord<Topology::DIM> CoordType;

    class SteererFeedback
    {
    public:
        SteererFeedback() : simulationEnd(false) {}

        void endSimulation() { simulationEnd = true; }
        bool simulationEnded() { return simulationEnd; }

    private:
        bool simulationEnd;
    };

    explicit Steerer(const unsigned period) : period(period) {}

    virtual ~Steerer() {}

    virtual Steerer *clone() const
    {
        throw std::logic_error("clone not implemented");
    }

    virtual void setRegion(const Region<Topology::DIM>& newRegion) { region = newRegion; }

    virtual void nextStep(
        GridType *grid,
        const Region<Topology::DIM>& validRegion,
        const CoordType& globalDimensions,
        unsigned step,
        SteererEvent event,
        std::size_t rank,
        bool lastCall,
        SteererFeedback *feedback) = 0;

    unsigned getPeriod() const { return period; }

protected:
    Region<Topology::DIM> region;
    unsigned period;
};

}

#endif
