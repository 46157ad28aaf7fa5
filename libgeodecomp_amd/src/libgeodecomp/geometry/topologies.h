// Boundary topologies: per axis either wrap-around (torus) or a constant
// edge cell outside the grid (cube).  Interface after
// src/libgeodecomp/geometry/topologies.h:176-332 (Topology::DIM,
// WrapsAxis<D>::VALUE, wrapsAxis(d), normalize(), isOutOfBounds(),
// Topologies::Cube<DIM>::Topology, Topologies::Torus<DIM>::Topology).
#ifndef LIBGEODECOMP_GEOMETRY_TOPOLOGIES_H
#define LIBGEODECOMP_GEOMETRY_TOPOLOGIES_H

#include <libgeodecomp/geometry/coord.h>

namespace LibGeoDecomp {

namespace TopologiesHelpers {

template<int DIMENSIONS, bool W0, bool W1, bool W2>
class Topology
{
public:
    static const int DIM = DIMENSIONS;

    template<int D>
    class WrapsAxis
    {
    public:
        static const bool VALUE = (D == 0) ? W0 : ((D == 1) ? W1 : W2);
    };

    __host__ __device__ static bool wrapsAxis(int d)
    {
        return d == 0 ? W0 : (d == 1 ? W1 : W2);
    }

    // coordinates on wrapping axes are folded into [0, dim); a coordinate
    // outside a non-wrapping axis is out of bounds
    __host__ __device__ static bool isOutOfBounds(const Coord<DIM>& c, const Coord<DIM>& dim)
    {
        for (int d = 0; d < DIM; ++d) {
            if (!wrapsAxis(d) && (c[d] < 0 || c[d] >= dim[d])) {
                return true;
            }
        }
        return false;
    }

    __host__ __device__ static Coord<DIM> normalize(const Coord<DIM>& c, const Coord<DIM>& dim)
    {
        Coord<DIM> r = c;
        for (int d = 0; d < DIM; ++d) {
            if (wrapsAxis(d) && dim[d] > 0) {
                r[d] = ((c[d] % dim[d]) + dim[d]) % dim[d];
            }
        }
        return r;
    }
};

}

class Topologies
{
public:
    template<int DIM>
    class Cube
    {
    public:
        typedef TopologiesHelpers::Topology<DIM, false, false, false> Topology;
    };

    template<int DIM>
    class Torus
    {
    public:
        typedef TopologiesHelpers::Topology<DIM, (DIM > 0), (DIM > 1), (DIM > 2)> Topology;
    };
};

}

#endif
