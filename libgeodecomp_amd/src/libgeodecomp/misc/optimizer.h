// Optimizer: searches a SimulationParameters space for the parameters an
// Evaluator rates highest (fitness: higher is better; simulation factories
// return minus the run time).  Interface after
// src/libgeodecomp/misc/optimizer.h:20-62.
#ifndef LIBGEODECOMP_MISC_OPTIMIZER_H
#define LIBGEODECOMP_MISC_OPTIMIZER_H

#include <libgeodecomp/misc/simulationparameters.h>

#include <limits>

namespace LibGeoDecomp {

class Optimizer
{
public:
    class Evaluator
    {
    public:
        virtual ~Evaluator() {}
        virtual double operator()(const SimulationParameters& params) = 0;
    };

    explicit Optimizer(SimulationParameters params) :
        params(params),
        fitness(std::numeric_limits<double>::lowest())
    {}

    virtual ~Optimizer() {}

    // at most maxSteps search steps; returns the best parameters found
    virtual SimulationParameters operator()(int maxSteps, Evaluator& eval) = 0;

    double getFitness() const { return fitness; }

protected:
    SimulationParameters params;
    double fitness;
};

}

#endif
