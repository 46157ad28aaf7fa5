// FixedCoord<X,Y,Z>: compile-time neighbour offset used by hood[FixedCoord<...>()]
// (src/libgeodecomp/geometry/fixedcoord.h).
#ifndef LIBGEODECOMP_GEOMETRY_FIXEDCOORD_H
#define LIBGEODECOMP_GEOMETRY_FIXEDCOORD_H

#include <libgeodecomp/misc/hostdevice.h>

namespace LibGeoDecomp {

template<int DIM_X = 0, int DIM_Y = 0, int DIM_Z = 0>
class FixedCoord
{
public:
    static const int X = DIM_X;
    static const int Y = DIM_Y;
    static const int Z = DIM_Z;

    __host__ __device__ inline FixedCoord() {}
};

}

#endif
