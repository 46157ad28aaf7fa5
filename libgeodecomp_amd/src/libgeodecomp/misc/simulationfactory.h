// SimulationFactory<CELL>: builds a simulator for a SimulationParameters
// point and rates it -- an Optimizer::Evaluator whose fitness is minus the
// wall time of run() (higher is better), and the worst possible fitness
// when the configuration fails to build or run.  Interface after
// src/libgeodecomp/misc/simulationfactory.h:21-131 and
// misc/cudasimulationfactory.h:21-71 (which maps a runtime error to the
// worst fitness); writers and steerers added to the factory are cloned into
// every simulator it builds.
#ifndef LIBGEODECOMP_MISC_SIMULATIONFACTORY_H
#define LIBGEODECOMP_MISC_SIMULATIONFACTORY_H

#include <libgeodecomp/io/clonableinitializer.h>
#include <libgeodecomp/misc/optimizer.h>
#include <libgeodecomp/parallelization/monolithicsimulator.h>

#include <chrono>
#include <limits>
#include <memory>
#include <stdexcept>
#include <vector>

namespace LibGeoDecomp {

template<typename CELL>
class SimulationFactory : public Optimizer::Evaluator
{
public:
    typedef std::shared_ptr<ClonableInitializer<CELL> > InitPtr;

    explicit SimulationFactory(InitPtr initializer) : initializer(initializer) {}
    virtual ~SimulationFactory() {}

    void addWriter(const Writer<CELL>& writer) { writers.push_back(std::shared_ptr<Writer<CELL> >(writer.clone())); }
    void addSteerer(const Steerer<CELL>& steerer)
    {
        steerers.push_back(std::shared_ptr<Steerer<CELL> >(steerer.clone()));
    }

    // a simulator with the factory's current parameters (caller owns it)
    Simulator<CELL> *operator()() { return buildSimulator(initializer, parameterSet); }

    virtual double operator()(const SimulationParameters& params)
    {
        try {
            std::unique_ptr<Simulator<CELL> > sim(buildSimulator(initializer, params));
            const auto t0 = std::chrono::steady_clock::now();
            sim->run();
            return -std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        } catch (const std::runtime_error&) {
            return std::numeric_limits<double>::lowest();
        } catch (const std::invalid_argument&) {
            return std::numeric_limits<double>::lowest();
        }
    }

    const SimulationParameters& parameters() const { return parameterSet; }
    SimulationParameters& parameters() { return parameterSet; }
    virtual std::string name() const = 0;

protected:
    InitPtr initializer;
    SimulationParameters parameterSet;
    std::vector<std::shared_ptr<Writer<CELL> > > writers;
    std::vector<std::shared_ptr<Steerer<CELL> > > steerers;

    virtual Simulator<CELL> *buildSimulator(InitPtr initializer, const SimulationParameters& params) const = 0;

    void addPlugins(MonolithicSimulator<CELL> *sim) const
    {
        for (std::size_t i = 0; i < writers.size(); ++i) {
            sim->addWriter(writers[i]->clone());
        }
        for (std::size_t i = 0; i < steerers.size(); ++i) {
            sim->addSteerer(steerers[i]->clone());
        }
    }
};

}

#endif
