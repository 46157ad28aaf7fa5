// PatternOptimizer: compass (pattern) search over SimulationParameters, the
// search of src/libgeodecomp/misc/patternoptimizer.cpp:11-170:
//  * start in the middle of every parameter's range with a step width of a
//    quarter of the range (or the granularity for tiny ranges), unless
//    step widths are given;
//  * each step evaluates the 2N neighbours middle +- width along every
//    parameter (skipping neighbours that clamping folds back onto the
//    middle and the point the search just came from) and moves to the best
//    one that is at least as fit as the best so far;
//  * when no neighbour is, all widths halve (not below the granularity);
//    the search ends when they all were at their minimum already.
#ifndef LIBGEODECOMP_MISC_PATTERNOPTIMIZER_H
#define LIBGEODECOMP_MISC_PATTERNOPTIMIZER_H

#include <libgeodecomp/misc/optimizer.h>

#include <stdexcept>
#include <vector>

namespace LibGeoDecomp {

class PatternOptimizer : public Optimizer
{
public:
    explicit PatternOptimizer(SimulationParameters params, std::vector<double> stepwidth = std::vector<double>(),
                              std::vector<double> minStepwidth = std::vector<double>()) :
        Optimizer(params),
        stepwidth(stepwidth),
        minStepwidth(minStepwidth)
    {
        if (stepwidth.empty() && minStepwidth.empty()) {
            for (std::size_t i = 0; i < this->params.size(); ++i) {
                const double range = this->params[i].getMax() - this->params[i].getMin();
                this->stepwidth.push_back(range / 4 >= 1 ? range / 4 : this->params[i].getGranularity());
                this->params[i].setValue(range / 2);
                this->minStepwidth.push_back(this->params[i].getGranularity());
            }
        }
        if (this->minStepwidth.empty() && this->stepwidth.size() == this->params.size()) {
            for (std::size_t i = 0; i < this->params.size(); ++i) {
                this->minStepwidth.push_back(this->params[i].getGranularity());
            }
        }
        if (this->stepwidth.size() != this->params.size() || this->minStepwidth.size() != this->params.size()) {
            throw std::invalid_argument("PatternOptimizer: one step width per parameter");
        }
    }

    SimulationParameters operator()(int maxSteps, Evaluator& eval)
    {
        SimulationParameters middle = params;
        fitness = eval(middle);
        std::size_t came = 0;   // pattern index of the last move (0: none)
        for (int k = 0; k < maxSteps; ++k) {
            const std::vector<SimulationParameters> pattern = neighbours(middle);
            const std::size_t best = bestNeighbour(pattern, eval, came);
            if (best == 0) {
                if (!halveWidths()) {
                    return middle;
                }
            } else {
                middle = pattern[best];
            }
            came = best;
        }
        return middle;
    }

    const std::vector<double>& stepwidths() const { return stepwidth; }

private:
    std::vector<double> stepwidth;
    std::vector<double> minStepwidth;

    // [middle, middle + w0 e0, middle - w0 e0, middle + w1 e1, ...]
    std::vector<SimulationParameters> neighbours(const SimulationParameters& middle) const
    {
        std::vector<SimulationParameters> out(1, middle);
        for (std::size_t i = 0; i < middle.size(); ++i) {
            SimulationParameters up(middle), down(middle);
            up[i] += stepwidth[i];
            down[i] += -stepwidth[i];
            out.push_back(up);
            out.push_back(down);
        }
        return out;
    }

    std::size_t bestNeighbour(const std::vector<SimulationParameters>& pattern, Evaluator& eval, std::size_t came)
    {
        std::size_t best = 0;
        for (std::size_t i = 1; i < pattern.size(); ++i) {
            const std::size_t dim = (i - 1) / 2;
            // clamped onto the middle: nothing new to evaluate
            if (pattern[0][dim].getValue() == pattern[i][dim].getValue()) {
                continue;
            }
            // the opposite of the last move leads back to where we were
            const std::size_t opposite = (i % 2) ? i + 1 : i - 1;
            if (came != 0 && opposite == came) {
                continue;
            }
            const double f = eval(pattern[i]);
            if (f >= fitness) {
                best = i;
                fitness = f;
            }
        }
        return best;
    }

    // false when every width was at its minimum already
    bool halveWidths()
    {
        bool changed = false;
        for (std::size_t i = 0; i < stepwidth.size(); ++i) {
            if (stepwidth[i] <= minStepwidth[i]) {
                stepwidth[i] = minStepwidth[i];
                continue;
            }
            stepwidth[i] = stepwidth[i] / 2 <= minStepwidth[i] ? minStepwidth[i] : stepwidth[i] / 2;
            changed = true;
        }
        return changed;
    }
};

}

#endif
