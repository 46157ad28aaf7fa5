// AutoTuningSimulator<CELL, OPTIMIZER>: tunes HipSimulator on short runs of
// the real problem, then runs the whole simulation with the best
// parameters found (src/libgeodecomp/parallelization/autotuningsimulator.h
// :59-289).  As there:
//  * the initializer is wrapped in a VarStepInitializerProxy; the number of
//    tuning steps is normalised so one tuning run takes about
//    `tuningSeconds` (the reference aims at one second: fitness goal -1);
//  * OPTIMIZER (e.g. PatternOptimizer) gets `optimizationSteps` steps over
//    the factory's parameters;
//  * writers and steerers added here are attached to the final run only
//    (tuning runs produce no output).
// The reference chooses between its serial, cache-blocking and CUDA
// factories; on one GPU the HIP factory is the only simulator.
#ifndef LIBGEODECOMP_PARALLELIZATION_AUTOTUNINGSIMULATOR_H
#define LIBGEODECOMP_PARALLELIZATION_AUTOTUNINGSIMULATOR_H

#include <libgeodecomp/io/varstepinitializerproxy.h>
#include <libgeodecomp/misc/hipsimulationfactory.h>

#include <algorithm>
#include <cmath>
#include <memory>

namespace LibGeoDecomp {

template<typename CELL, typename OPTIMIZER>
class AutoTuningSimulator
{
public:
    explicit AutoTuningSimulator(Initializer<CELL> *initializer, unsigned optimizationSteps = 10,
                                 double tuningSeconds = 0.05) :
        optimizationSteps(optimizationSteps),
        tuningSeconds(tuningSeconds),
        proxy(new VarStepInitializerProxy<CELL>(initializer)),
        factory(proxy),
        tuningSteps(0),
        bestFitness(0)
    {}

    void addWriter(Writer<CELL> *writer) { writers.push_back(std::shared_ptr<Writer<CELL> >(writer)); }
    void addSteerer(Steerer<CELL> *steerer) { steerers.push_back(std::shared_ptr<Steerer<CELL> >(steerer)); }

    void run()
    {
        const unsigned total = proxy->maxSteps() - proxy->startStep();
        tuningSteps = normalizeSteps(total);
        proxy->setMaxSteps(tuningSteps);
        OPTIMIZER optimizer(factory.parameters());
        best = optimizer(optimizationSteps, factory);
        bestFitness = optimizer.getFitness();
        proxy->resetMaxSteps();
        const HipKernels::LaunchChoice c = HipSimulationFactory<CELL>::choice(best);
        HipSimulator<CELL> sim(proxy->clone(), c.blockSize, c.stepsPerLaunch, c.variant);
        for (std::size_t i = 0; i < writers.size(); ++i) {
            sim.addWriter(writers[i]->clone());
        }
        for (std::size_t i = 0; i < steerers.size(); ++i) {
            sim.addSteerer(steerers[i]->clone());
        }
        sim.run();
        finalStep = sim.getStep();
        finalGrid.reset(new Grid<CELL, typename HipSimulator<CELL>::Topology>(proxy->gridBox()));
        const GridBase<CELL, HipSimulator<CELL>::DIM> *g = sim.getGrid();
        CoordBox<HipSimulator<CELL>::DIM> box = g->boundingBox();
        for (typename CoordBox<HipSimulator<CELL>::DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
            finalGrid->set(*i, g->get(*i));
        }
        finalGrid->setEdge(g->getEdge());
    }

    const SimulationParameters& bestParameters() const { return best; }
    double bestTuningFitness() const { return bestFitness; }
    unsigned stepsPerTuningRun() const { return tuningSteps; }
    unsigned getStep() const { return finalStep; }
    // the final run's grid (valid after run())
    const GridBase<CELL, HipSimulator<CELL>::DIM> *getGrid() const { return finalGrid.get(); }

private:
    unsigned optimizationSteps;
    double tuningSeconds;
    std::shared_ptr<VarStepInitializerProxy<CELL> > proxy;
    HipSimulationFactory<CELL> factory;
    SimulationParameters best;
    unsigned tuningSteps;
    double bestFitness;
    unsigned finalStep = 0;
    std::unique_ptr<Grid<CELL, typename HipSimulator<CELL>::Topology> > finalGrid;
    std::vector<std::shared_ptr<Writer<CELL> > > writers;
    std::vector<std::shared_ptr<Steerer<CELL> > > steerers;

    // steps for which one run of the default configuration takes about
    // tuningSeconds (at least 1, at most the whole run)
    unsigned normalizeSteps(unsigned total)
    {
        unsigned steps = std::min(total, 5u);
        if (steps == 0) {
            return 0;
        }
        for (int round = 0; round < 4; ++round) {
            proxy->setMaxSteps(steps);
            const double seconds = -factory(factory.parameters());
            if (!(seconds > 0)) {
                break;
            }
            const unsigned next = std::max(1u, std::min(total, (unsigned)std::lround(steps * tuningSeconds / seconds)));
            if (next == steps) {
                break;
            }
            steps = next;
        }
        return steps;
    }
};

}

#endif
