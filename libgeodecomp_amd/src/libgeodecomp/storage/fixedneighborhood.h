// FixedNeighborhood<CELL>: the neighbourhood a struct-of-arrays cell's
// static updateLineX(hoodOld, indexEnd, hoodNew, nanoStep) reads through,
// on the host (the oracle's SoA update) and in HIP kernels (the SoA kernel
// family).  Interface and boundary semantics after
// src/libgeodecomp/storage/fixedneighborhood.h:16-120:
//   hood[FixedCoord<X, Y, Z>()]  accessor of the cell at the current index
//                                plus (X, Y, Z), where a negative X adds
//                                offsetWest, a positive X offsetEast (and
//                                likewise Top/Bottom for Y, South/North
//                                for Z) to the coordinate -- the wrap of a
//                                periodic axis, nonzero only for the first
//                                or last cell of that axis;
//   hood.index(), hood += n      the position shared with the accessor it
//                                was built on, so the cell's loop advances
//                                both.
// The neighbourhood holds a reference to that accessor, as the reference's
// does, so a streak split into peeled boundary cells and a middle part
// continues from one index.
#ifndef LIBGEODECOMP_STORAGE_FIXEDNEIGHBORHOOD_H
#define LIBGEODECOMP_STORAGE_FIXEDNEIGHBORHOOD_H

#include <libflatarray/flat_array.hpp>
#include <libgeodecomp/geometry/fixedcoord.h>

namespace LibGeoDecomp {

template<typename CELL>
class FixedNeighborhood
{
public:
    typedef CELL Cell;
    typedef LibFlatArray::soa_accessor<CELL> Accessor;

    __host__ __device__ inline FixedNeighborhood(Accessor& accessor, long offsetWest = 0, long offsetEast = 0,
                                                 long offsetTop = 0, long offsetBottom = 0, long offsetSouth = 0,
                                                 long offsetNorth = 0) :
        accessor(accessor),
        offsetWest(offsetWest),
        offsetEast(offsetEast),
        offsetTop(offsetTop),
        offsetBottom(offsetBottom),
        offsetSouth(offsetSouth),
        offsetNorth(offsetNorth)
    {}

    template<int X, int Y, int Z>
    __host__ __device__ inline Accessor operator[](FixedCoord<X, Y, Z>) const
    {
        const long x = X + (X < 0 ? offsetWest : 0) + (X > 0 ? offsetEast : 0);
        const long y = Y + (Y < 0 ? offsetTop : 0) + (Y > 0 ? offsetBottom : 0);
        const long z = Z + (Z < 0 ? offsetSouth : 0) + (Z > 0 ? offsetNorth : 0);
        return Accessor(accessor.data(), accessor.index() + accessor.gen_index(x, y, z), accessor.dim_x(),
                        accessor.dim_xy(), accessor.dim_prod());
    }

    __host__ __device__ inline void operator>>(CELL& cell) const { accessor >> cell; }

    __host__ __device__ inline long& index() { return accessor.index(); }
    __host__ __device__ inline const long& index() const { return accessor.index(); }
    __host__ __device__ inline void operator+=(long offset) { accessor += offset; }

private:
    Accessor& accessor;
    long offsetWest;
    long offsetEast;
    long offsetTop;
    long offsetBottom;
    long offsetSouth;
    long offsetNorth;
};

}

#endif
