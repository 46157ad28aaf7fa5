// Streak<DIM>: a run of cells along x, [origin.x, endX)
// (src/libgeodecomp/geometry/streak.h).
#ifndef LIBGEODECOMP_GEOMETRY_STREAK_H
#define LIBGEODECOMP_GEOMETRY_STREAK_H

#include <libgeodecomp/geometry/coord.h>

namespace LibGeoDecomp {

template<int DIM>
class Streak
{
public:
    Coord<DIM> origin;
    int endX;

    explicit Streak(const Coord<DIM>& origin = Coord<DIM>(), int endX = 0) :
        origin(origin),
        endX(endX)
    {}

    int length() const { return endX - origin.x(); }
    bool operator==(const Streak& o) const { return origin == o.origin && endX == o.endX; }
};

}

#endif
