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
// Beyond the reference: with setCacheFile(path), the best parameters are
// kept in a text file keyed by the cell type, the grid box, the tuning
// space's parameter names and the device name; a later run with the same
// key takes them from the file and skips the search (one line per key,
// "key<TAB>index index ...", the last line of a key wins).
// The reference chooses between its serial, cache-blocking and CUDA
// factories; on one GPU the HIP factory is the only simulator.
#ifndef LIBGEODECOMP_PARALLELIZATION_AUTOTUNINGSIMULATOR_H
#define LIBGEODECOMP_PARALLELIZATION_AUTOTUNINGSIMULATOR_H

#include <libgeodecomp/io/varstepinitializerproxy.h>
#include <libgeodecomp/misc/hipsimulationfactory.h>

#include <algorithm>
#include <cmath>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <typeinfo>

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

    // persist the tuning result across runs (see above); "" turns it off
    void setCacheFile(const std::string& path) { cacheFile = path; }
    // whether the last run() took its parameters from the cache file
    bool usedCache() const { return fromCache; }

    void run()
    {
        fromCache = !cacheFile.empty() && readCache(&best);
        if (fromCache) {
            tuningSteps = 0;
            bestFitness = 0;
        } else {
            const unsigned total = proxy->maxSteps() - proxy->startStep();
            tuningSteps = normalizeSteps(total);
            proxy->setMaxSteps(tuningSteps);
            OPTIMIZER optimizer(factory.parameters());
            best = optimizer(optimizationSteps, factory);
            bestFitness = optimizer.getFitness();
            proxy->resetMaxSteps();
            if (!cacheFile.empty()) {
                writeCache(best);
            }
        }
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
    std::string cacheFile;
    bool fromCache = false;
    std::unique_ptr<Grid<CELL, typename HipSimulator<CELL>::Topology> > finalGrid;
    std::vector<std::shared_ptr<Writer<CELL> > > writers;
    std::vector<std::shared_ptr<Steerer<CELL> > > steerers;

    // cell type, grid box, parameter names, device: what a tuning result
    // depends on (no whitespace inside, so it is one token)
    std::string cacheKey()
    {
        std::ostringstream k;
        k << typeid(CELL).name() << '|';
        const CoordBox<HipSimulator<CELL>::DIM> box = proxy->gridBox();
        for (int d = 0; d < HipSimulator<CELL>::DIM; ++d) {
            k << box.dimensions[d] << (d + 1 < HipSimulator<CELL>::DIM ? "x" : "|");
        }
        const SimulationParameters& p = factory.parameters();
        for (std::size_t i = 0; i < p.size(); ++i) {
            k << p.name(i) << ',';
        }
        int dev = 0;
        char name[256] = "unknown";
        if (lgdhip_get_device(&dev) == 0) {
            lgdhip_device_name(dev, name, (int)sizeof(name));
        }
        std::string device(name);
        for (char& c : device) {
            if (c == ' ' || c == '\t') {
                c = '_';
            }
        }
        k << '|' << device;
        return k.str();
    }

    // the parameters stored for this key (their index values), if any
    bool readCache(SimulationParameters *out)
    {
        std::ifstream in(cacheFile.c_str());
        if (!in) {
            return false;
        }
        const std::string key = cacheKey();
        std::string line, found;
        while (std::getline(in, line)) {
            if (line.compare(0, key.size() + 1, key + "\t") == 0) {
                found = line.substr(key.size() + 1);
            }
        }
        if (found.empty()) {
            return false;
        }
        SimulationParameters p = factory.parameters();
        std::istringstream values(found);
        for (std::size_t i = 0; i < p.size(); ++i) {
            double v;
            if (!(values >> v) || v < p[i].getMin() || v >= p[i].getMax()) {
                return false;  // stale or damaged line: tune again
            }
            p[i].setValue(v);
        }
        *out = p;
        return true;
    }

    void writeCache(const SimulationParameters& p)
    {
        std::ofstream out(cacheFile.c_str(), std::ios::app);
        out << cacheKey() << '\t';
        for (std::size_t i = 0; i < p.size(); ++i) {
            out << p[i].getValue() << (i + 1 < p.size() ? " " : "\n");
        }
    }

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
