// VarStepInitializerProxy<CELL>: wraps an Initializer and overrides its step
// count, so the autotuner can time short runs of the real problem and then
// run it to the end (src/libgeodecomp/io/varstepinitializerproxy.h:20-110).
// Clones share the wrapped initializer; maxSteps() = startStep() + the
// number of steps set (initially the wrapped run's).
#ifndef LIBGEODECOMP_IO_VARSTEPINITIALIZERPROXY_H
#define LIBGEODECOMP_IO_VARSTEPINITIALIZERPROXY_H

#include <libgeodecomp/io/clonableinitializer.h>

#include <memory>

namespace LibGeoDecomp {

template<typename CELL>
class VarStepInitializerProxy : public ClonableInitializer<CELL>
{
public:
    typedef typename Initializer<CELL>::Topology Topology;
    static const int DIM = Topology::DIM;

    explicit VarStepInitializerProxy(Initializer<CELL> *wrapped) :
        wrapped(wrapped),
        steps(wrapped->maxSteps() - wrapped->startStep())
    {}

    void setMaxSteps(unsigned n) { steps = n; }
    void resetMaxSteps() { steps = wrapped->maxSteps() - wrapped->startStep(); }

    virtual void grid(GridBase<CELL, DIM> *target) { wrapped->grid(target); }
    virtual Coord<DIM> gridDimensions() const { return wrapped->gridDimensions(); }
    virtual CoordBox<DIM> gridBox() { return wrapped->gridBox(); }
    virtual unsigned startStep() const { return wrapped->startStep(); }
    virtual unsigned maxSteps() const { return wrapped->startStep() + steps; }
    virtual ClonableInitializer<CELL> *clone() const { return new VarStepInitializerProxy(*this); }

private:
    std::shared_ptr<Initializer<CELL> > wrapped;
    unsigned steps;
};

}

#endif
