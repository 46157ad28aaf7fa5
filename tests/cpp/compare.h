// Helpers of the GPU-vs-oracle C++ tests: run one simulation through
// HipSimulator and the oracle SerialSimulator with recording writers (and
// optionally steerers) and compare the grids at every writer callback and
// the event sequences; probe-cell checks.
#ifndef LGD_TEST_COMPARE_H
#define LGD_TEST_COMPARE_H

#include "devicecells.h"
#include "harness.h"

#include <libgeodecomp/parallelization/hipsimulator.h>
#include <serialsimulator.h>

#include <cstdio>
#include <functional>
#include <vector>

template<typename CELL>
struct Snapshots
{
    typedef typename Writer<CELL>::GridType GridType;
    std::vector<std::vector<CELL> > grids;

    static void hook(const GridType& g, unsigned, void *ctx)
    {
        Snapshots *s = static_cast<Snapshots *>(ctx);
        std::vector<CELL> cells;
        CoordBox<GridType::DIM> box = g.boundingBox();
        for (typename CoordBox<GridType::DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
            cells.push_back(g.get(*i));
        }
        s->grids.push_back(cells);
    }
};

// runs INIT's simulation through both simulators with an event-recording
// writer of the given period; returns whether grids and events agree
template<typename CELL>
bool runBoth(std::function<Initializer<CELL> *()> make, unsigned writerPeriod, Coord<3> block,
             Snapshots<CELL> *snapHip, Snapshots<CELL> *snapRef, unsigned steererPeriod = 0,
             typename RecordingSteerer<CELL>::Hook steer = 0)
{
    std::vector<Event> evHip, evRef;
    HipSimulator<CELL> hip(make(), block);
    SerialSimulator<CELL> ref(make());
    hip.addWriter(new RecordingWriter<CELL>(&evHip, writerPeriod, &Snapshots<CELL>::hook, snapHip));
    ref.addWriter(new RecordingWriter<CELL>(&evRef, writerPeriod, &Snapshots<CELL>::hook, snapRef));
    if (steererPeriod) {
        hip.addSteerer(new RecordingSteerer<CELL>(&evHip, steererPeriod, steer));
        ref.addSteerer(new RecordingSteerer<CELL>(&evRef, steererPeriod, steer));
    }
    hip.run();
    ref.run();
    CHECK_EQ(hip.getStep(), ref.getStep());
    CHECK(evHip == evRef);
    CHECK_EQ(snapHip->grids.size(), snapRef->grids.size());
    const bool same = snapHip->grids == snapRef->grids;
    CHECK(same);
    return same;
}

// every cell of the last snapshot valid, with cycle == steps * nano steps
template<typename CELL>
void checkProbes(const Snapshots<CELL>& s, unsigned steps, unsigned nano)
{
    CHECK(!s.grids.empty());
    if (s.grids.empty()) {
        return;
    }
    int invalid = 0, badBits = 0, wrongCycle = 0;
    for (const CELL& c : s.grids.back()) {
        invalid += !c.valid;
        badBits |= c.bad;
        wrongCycle += c.cycle != steps * nano;
    }
    if (invalid || wrongCycle) {
        std::printf("  %d invalid cells (failure bits 0x%x), %d with a wrong cycle count\n", invalid, badBits,
                    wrongCycle);
    }
    CHECK_EQ(invalid, 0);
    CHECK_EQ(wrongCycle, 0);
}

template<typename CELL>
void probeCase(const Coord<CELL::API::Topology::DIM>& dim, unsigned steps, unsigned writerPeriod,
               Coord<3> block = Coord<3>())
{
    typedef ProbeInitializer<CELL> Init;
    Snapshots<CELL> sh, sr;
    runBoth<CELL>([&]() { return new Init(dim, steps, 20261202); }, writerPeriod, block, &sh, &sr);
    checkProbes(sh, steps, APITraits::SelectNanoSteps<CELL>::VALUE);
    checkProbes(sr, steps, APITraits::SelectNanoSteps<CELL>::VALUE);
}

#endif
