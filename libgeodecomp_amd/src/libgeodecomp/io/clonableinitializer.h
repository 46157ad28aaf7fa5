// ClonableInitializer<CELL>: an Initializer that can copy itself, so a
// simulation factory can build one simulator per evaluation
// (src/libgeodecomp/io/clonableinitializer.h:12-24).
#ifndef LIBGEODECOMP_IO_CLONABLEINITIALIZER_H
#define LIBGEODECOMP_IO_CLONABLEINITIALIZER_H

#include <libgeodecomp/io/initializer.h>

namespace LibGeoDecomp {

template<typename CELL>
class ClonableInitializer : public Initializer<CELL>
{
public:
    typedef void IsClonable;
    virtual ClonableInitializer<CELL> *clone() const = 0;
};

}

#endif
