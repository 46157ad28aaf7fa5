// APITraits: a Cell declares its capabilities by deriving its nested
// `class API` from APITraits::HasFoo; simulators query them through
// APITraits::SelectFoo<CELL>.  Same trait names and detection scheme as
// src/libgeodecomp/misc/apitraits.h (Has*/Select* pairs, :150-364, :813-836):
// the trait classes export a `SupportsFoo` typedef that a partial
// specialisation of SelectFoo detects.
//
// Added for the HIP path: HasHipKernel<FAMILY> / SelectHipKernel<CELL>,
// naming the hand-written kernel family that implements the cell's update
// (see libgeodecomp/hip/kernelfamilies.h).  It can also be attached after
// the class with LIBGEODECOMP_REGISTER_HIP_KERNEL.  Cells without a family
// run through the generic family (libgeodecomp/hip/genericfamily.h), which
// calls update(), or updateHIP() / updateCUDA() under
// HasSeparateHIPUpdate / HasSeparateCUDAUpdate.
#ifndef LIBGEODECOMP_MISC_APITRAITS_H
#define LIBGEODECOMP_MISC_APITRAITS_H

#include <libgeodecomp/geometry/stencils.h>
#include <libgeodecomp/geometry/topologies.h>

namespace LibGeoDecomp {

class APITraits
{
public:
    class TrueType {};
    class FalseType {};

    template<typename T>
    class AlwaysVoid
    {
    public:
        typedef void Type;
    };

    template<typename T, typename DUMMY = void>
    class IsTrue { public: static const bool VALUE = false; };
    template<typename DUMMY>
    class IsTrue<TrueType, DUMMY> { public: static const bool VALUE = true; };

    // ---- struct-of-arrays storage
    template<typename CELL, typename HAS = void>
    class SelectSoA { public: typedef FalseType Value; };
    template<typename CELL>
    class SelectSoA<CELL, typename CELL::API::SupportsSoA> { public: typedef TrueType Value; };
    class HasSoA { public: typedef void SupportsSoA; };

    // ---- line-wise update
    template<typename CELL, typename HAS = void>
    class SelectUpdateLineX { public: typedef FalseType Value; };
    template<typename CELL>
    class SelectUpdateLineX<CELL, typename CELL::API::SupportsUpdateLineX> { public: typedef TrueType Value; };
    class HasUpdateLineX { public: typedef void SupportsUpdateLineX; };

    // ---- neighbourhood accessed with FixedCoord only
    template<typename CELL, typename HAS = void>
    class SelectFixedCoordsOnlyUpdate { public: typedef FalseType Value; };
    template<typename CELL>
    class SelectFixedCoordsOnlyUpdate<CELL, typename CELL::API::SupportsFixedCoordsOnlyUpdate>
    { public: typedef TrueType Value; };
    class HasFixedCoordsOnlyUpdate { public: typedef void SupportsFixedCoordsOnlyUpdate; };

    // ---- stencil (default Moore<2,1>)
    template<typename CELL, typename HAS = void>
    class SelectStencil { public: typedef Stencils::Moore<2, 1> Value; };
    template<typename CELL>
    class SelectStencil<CELL, typename CELL::API::SupportsStencil> { public: typedef typename CELL::API::Stencil Value; };
    template<typename STENCIL>
    class HasStencil { public: typedef void SupportsStencil; typedef STENCIL Stencil; };

    // ---- nano steps (default 1)
    template<typename CELL, typename HAS = void>
    class SelectNanoSteps { public: static const unsigned VALUE = 1; };
    template<typename CELL>
    class SelectNanoSteps<CELL, typename CELL::API::SupportsNanoSteps>
    { public: static const unsigned VALUE = CELL::API::NANO_STEPS; };
    template<unsigned CELL_NANO_STEPS>
    class HasNanoSteps { public: typedef void SupportsNanoSteps; static const unsigned NANO_STEPS = CELL_NANO_STEPS; };

    // ---- topology (default Cube<2>)
    template<typename CELL, typename HAS = void>
    class SelectTopology { public: typedef Topologies::Cube<2>::Topology Value; };
    template<typename CELL>
    class SelectTopology<CELL, typename CELL::API::SupportsTopology> { public: typedef typename CELL::API::Topology Value; };
    template<typename TOPOLOGY>
    class HasTopology { public: typedef void SupportsTopology; typedef TOPOLOGY Topology; };
    template<int DIM>
    class HasCubeTopology : public HasTopology<typename Topologies::Cube<DIM>::Topology> {};
    template<int DIM>
    class HasTorusTopology : public HasTopology<typename Topologies::Torus<DIM>::Topology> {};

    // ---- separate device update (the reference's updateCUDA; here updateHIP)
    template<typename CELL, typename HAS = void>
    class SelectSeparateCUDAUpdate { public: typedef FalseType Value; };
    template<typename CELL>
    class SelectSeparateCUDAUpdate<CELL, typename CELL::API::SupportsSeparateCUDAUpdate>
    { public: typedef TrueType Value; };
    class HasSeparateCUDAUpdate { public: typedef void SupportsSeparateCUDAUpdate; };
    // the HIP spelling: the generic kernel family calls updateHIP()
    template<typename CELL, typename HAS = void>
    class SelectSeparateHIPUpdate { public: typedef FalseType Value; };
    template<typename CELL>
    class SelectSeparateHIPUpdate<CELL, typename CELL::API::SupportsSeparateHIPUpdate>
    { public: typedef TrueType Value; };
    class HasSeparateHIPUpdate { public: typedef void SupportsSeparateHIPUpdate; };

    // ---- static data block (default char)
    template<typename CELL, typename HAS = void>
    class SelectStaticData { public: typedef char Value; };
    template<typename CELL>
    class SelectStaticData<CELL, typename CELL::API::SupportsStaticData>
    { public: typedef typename CELL::API::StaticData Value; };
    template<typename DATA>
    class HasStaticData { public: typedef void SupportsStaticData; typedef DATA StaticData; };

    // ---- the update path the reference's UpdateFunctor selector picks
    // (storage/updatefunctor.h:96-139, :140-174): a struct-of-arrays cell with
    // a FixedCoord-only neighbourhood and a line update runs streak by streak
    // through static updateLineX(hoodOld, indexEnd, hoodNew, nanoStep); an
    // array-of-structs cell with both of those traits runs static
    // updateLineX(target, &x, endX, hood, nanoStep) over line pointers; every
    // other cell runs update(hood, nanoStep) cell by cell.
    template<typename CELL>
    class SelectSoAStreakUpdate
    {
    public:
        static const bool VALUE = IsTrue<typename SelectSoA<CELL>::Value>::VALUE &&
                                  IsTrue<typename SelectUpdateLineX<CELL>::Value>::VALUE &&
                                  IsTrue<typename SelectFixedCoordsOnlyUpdate<CELL>::Value>::VALUE;
    };
    template<typename CELL>
    class SelectLinePointerUpdate
    {
    public:
        static const bool VALUE = !IsTrue<typename SelectSoA<CELL>::Value>::VALUE &&
                                  IsTrue<typename SelectUpdateLineX<CELL>::Value>::VALUE &&
                                  IsTrue<typename SelectFixedCoordsOnlyUpdate<CELL>::Value>::VALUE;
    };

    // ---- hand-written HIP kernel family implementing the cell's update
    template<typename CELL, typename HAS = void>
    class SelectHipKernel { public: typedef void Value; };
    template<typename CELL>
    class SelectHipKernel<CELL, typename CELL::API::SupportsHipKernel>
    { public: typedef typename CELL::API::HipKernel Value; };
    template<typename FAMILY>
    class HasHipKernel { public: typedef void SupportsHipKernel; typedef FAMILY HipKernel; };
};

}

// Names the kernel family of CELL after its definition (member pointers of
// CELL are only usable once the class is complete), e.g.
//   LIBGEODECOMP_REGISTER_HIP_KERNEL(ConwayCell, LibGeoDecomp::HipKernels::Life2D<ConwayCell, bool, &ConwayCell::alive>)
#define LIBGEODECOMP_REGISTER_HIP_KERNEL(CELL, ...)                           \
    namespace LibGeoDecomp {                                                    \
    template<>                                                                  \
    class APITraits::SelectHipKernel<CELL, void>                                \
    {                                                                           \
    public:                                                                     \
        typedef __VA_ARGS__ Value;                                              \
    };                                                                          \
    }

#endif
