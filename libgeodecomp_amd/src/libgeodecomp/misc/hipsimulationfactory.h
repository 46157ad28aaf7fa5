// HipSimulationFactory<CELL>: the simulation factory of HipSimulator -- the
// counterpart of src/libgeodecomp/misc/cudasimulationfactory.h:21-71.  Its
// parameters are the tuning space of the cell's kernel family (the
// reference's three block dimensions become what each family can actually
// vary: work-group shape, temporal-blocking depth, kernel variant):
//   Jacobi3DF64   Variant in the order 1 2 3 13 11 12 7 4 8 5 6 9 10: single
//                 sweeps (plane marches, one and two cells per thread, then
//                 the tile pipeline at K = 1
//                 with one and two rows per thread), temporal blocking of
//                 growing depth (K = 2, 2, 3, 3, 4), then K = 4 with two rows
//                 per thread (16-, 32-row tiles);
//                 ZChunk {0, 8, 16, 26, 32, 48, 64}: z planes per work-group
//                 (0: the launch cost model's choice, the starting point)
//   Life2D        StepsPerLaunch [1, 33), Variant {1..8}
//   LBMD3Q19F32   BlockDimY {1, 2, 4, 8}, Variant {0, 1}
//   Generic       BlockDimX {64, 128, 256}, BlockDimY {1, 2, 4, 8}
// Discrete values are ordered so that neighbours in the order are similar
// configurations (the optimizers step through indices).  A configuration
// that a family rejects rates the worst fitness.
#ifndef LIBGEODECOMP_MISC_HIPSIMULATIONFACTORY_H
#define LIBGEODECOMP_MISC_HIPSIMULATIONFACTORY_H

#include <libgeodecomp/misc/simulationfactory.h>
#include <libgeodecomp/parallelization/hipsimulator.h>

namespace LibGeoDecomp {

namespace HipKernels {

// HipSimulator constructor arguments of one tuning point
struct LaunchChoice
{
    Coord<3> blockSize;
    int stepsPerLaunch;
    int variant;
    LaunchChoice() : blockSize(0, 0, 0), stepsPerLaunch(0), variant(0) {}
};

template<typename FAMILY>
class TuningSpace
{
public:
    static void add(SimulationParameters *p)
    {
        p->addParameter("BlockDimX", std::vector<int>{64, 128, 256});
        p->addParameter("BlockDimY", std::vector<int>{1, 2, 4, 8});
    }
    static LaunchChoice choose(const SimulationParameters& p)
    {
        LaunchChoice c;
        c.blockSize = Coord<3>((int)p["BlockDimX"], (int)p["BlockDimY"], 1);
        return c;
    }
};

template<typename CELL, typename VALUE, VALUE CELL::*MEMBER>
class TuningSpace<Jacobi3DF64<CELL, VALUE, MEMBER> >
{
public:
    static void add(SimulationParameters *p)
    {
        p->addParameter("Variant", std::vector<int>{1, 2, 3, 13, 11, 12, 7, 4, 8, 5, 6, 9, 10});
        p->addParameter("ZChunk", std::vector<int>{0, 8, 16, 26, 32, 48, 64});
    }
    static LaunchChoice choose(const SimulationParameters& p)
    {
        LaunchChoice c;
        c.variant = (int)p["Variant"];
        c.blockSize = Coord<3>(0, 0, (int)p["ZChunk"]);
        return c;
    }
};

template<typename CELL, typename VALUE, VALUE CELL::*MEMBER>
class TuningSpace<Life2D<CELL, VALUE, MEMBER> >
{
public:
    static void add(SimulationParameters *p)
    {
        p->addParameter("StepsPerLaunch", 1, 33);
        p->addParameter("Variant", std::vector<int>{1, 2, 3, 4, 5, 6, 7, 8});
    }
    static LaunchChoice choose(const SimulationParameters& p)
    {
        LaunchChoice c;
        c.stepsPerLaunch = (int)p["StepsPerLaunch"];
        c.variant = (int)p["Variant"];
        return c;
    }
};

template<typename CELL>
class TuningSpace<LBMD3Q19F32<CELL> >
{
public:
    static void add(SimulationParameters *p)
    {
        p->addParameter("BlockDimY", std::vector<int>{1, 2, 4, 8});
        p->addParameter("Variant", std::vector<int>{0, 1});
    }
    static LaunchChoice choose(const SimulationParameters& p)
    {
        LaunchChoice c;
        c.blockSize = Coord<3>(0, (int)p["BlockDimY"], 0);
        c.variant = (int)p["Variant"];
        return c;
    }
};

}

template<typename CELL>
class HipSimulationFactory : public SimulationFactory<CELL>
{
public:
    typedef typename SimulationFactory<CELL>::InitPtr InitPtr;
    typedef typename HipSimulator<CELL>::Family Family;

    explicit HipSimulationFactory(InitPtr initializer) : SimulationFactory<CELL>(initializer)
    {
        HipKernels::TuningSpace<Family>::add(&this->parameterSet);
    }

    std::string name() const { return "HipSimulator"; }

    // the constructor arguments a parameter point maps to
    static HipKernels::LaunchChoice choice(const SimulationParameters& params)
    {
        return HipKernels::TuningSpace<Family>::choose(params);
    }

protected:
    Simulator<CELL> *buildSimulator(InitPtr initializer, const SimulationParameters& params) const
    {
        const HipKernels::LaunchChoice c = choice(params);
        HipSimulator<CELL> *sim = new HipSimulator<CELL>(initializer->clone(), c.blockSize, c.stepsPerLaunch,
                                                         c.variant);
        this->addPlugins(sim);
        return sim;
    }
};

}

#endif
