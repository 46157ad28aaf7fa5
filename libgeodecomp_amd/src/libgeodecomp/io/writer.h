// Writer<CELL>: output plugin called back by a MonolithicSimulator
// (src/libgeodecomp/io/writer.h:29-111): WRITER_INITIALIZED once,
// WRITER_STEP_FINISHED when step % period == 0, WRITER_ALL_DONE at the
// last step; a zero period is rejected with std::invalid_argument.
#ifndef LIBGEODECOMP_IO_WRITER_H
#define LIBGEODECOMP_IO_WRITER_H

#include <libgeodecomp/misc/apitraits.h>
#include <libgeodecomp/storage/gridbase.h>

#include <stdexcept>
#include <string>

namespace LibGeoDecomp {

enum WriterEvent {
    WRITER_INITIALIZED,
    WRITER_STEP_FINISHED,
    WRITER_ALL_DONE
};

template<typename CELL_TYPE>
class Writer
{
public:
    typedef typename APITraits::SelectTopology<CELL_TYPE>::Value Topology;
    const static int DIM = Topology::DIM;
    typedef GridBase<CELL_TYPE, DIM> GridType;
    static const unsigned NANO_STEPS = APITraits::SelectNanoSteps<CELL_TYPE>::VALUE;

    Writer(const std::string& prefix, const unsigned period) :
        prefix(prefix),
        period(period)
    {
        if (period == 0) {
            throw std::invalid_argument("period must be positive");
        }
    }

    virtual ~Writer() {}

    virtual Writer *clone() const = 0;

    virtual void stepFinished(const GridType& grid, unsigned step, WriterEvent event) = 0;

    unsigned getPeriod() const { return period; }
    const std::string& getPrefix() const { return prefix; }

protected:
    std::string prefix;
    unsigned period;
};

}

#endif
