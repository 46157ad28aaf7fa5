// Stencil shapes (src/libgeodecomp/geometry/stencils.h:176-251): Moore
// (Chebyshev ball) and VonNeumann (Manhattan ball) with RADIUS and VOLUME.
#ifndef LIBGEODECOMP_GEOMETRY_STENCILS_H
#define LIBGEODECOMP_GEOMETRY_STENCILS_H

namespace LibGeoDecomp {

namespace StencilsHelpers {

template<int BASE, int EXP>
class Power
{
public:
    static const int VALUE = BASE * Power<BASE, EXP - 1>::VALUE;
};

template<int BASE>
class Power<BASE, 0>
{
public:
    static const int VALUE = 1;
};

// number of lattice points with |x|_1 <= R in DIM dimensions:
// V(D, R) = V(D-1, R) + 2 * sum_{r=0}^{R-1} V(D-1, r)
template<int DIM, int R>
class VonNeumannVolume;

template<int DIM, int R>
class SumLower
{
public:
    static const int VALUE = VonNeumannVolume<DIM - 1, R>::VALUE + SumLower<DIM, R - 1>::VALUE;
};

template<int DIM>
class SumLower<DIM, -1>
{
public:
    static const int VALUE = 0;
};

template<int DIM, int R>
class VonNeumannVolume
{
public:
    static const int VALUE = VonNeumannVolume<DIM - 1, R>::VALUE + 2 * SumLower<DIM, R - 1>::VALUE;
};

template<int R>
class VonNeumannVolume<0, R>
{
public:
    static const int VALUE = 1;
};

}

class Stencils
{
public:
    template<int DIMENSIONS, int RAD>
    class Moore
    {
    public:
        static const int DIM = DIMENSIONS;
        static const int RADIUS = RAD;
        static const int VOLUME = StencilsHelpers::Power<2 * RAD + 1, DIMENSIONS>::VALUE;
    };

    template<int DIMENSIONS, int RAD>
    class VonNeumann
    {
    public:
        static const int DIM = DIMENSIONS;
        static const int RADIUS = RAD;
        static const int VOLUME = StencilsHelpers::VonNeumannVolume<DIMENSIONS, RAD>::VALUE;
    };
};

}

#endif
