// SerialSimulator<CELL> -- CPU oracle of the simulator semantics.
// TEST INFRASTRUCTURE ONLY: C++ tests run a user Cell through this class and
// through HipSimulator and compare grids and plugin event sequences.  No
// product code includes it.
//
// Restates src/libgeodecomp/parallelization/serialsimulator.h:
//  * constructor initialises both grids from the Initializer (:48-59);
//  * step(): steerers (STEERER_NEXT_STEP when step % period == 0), all nano
//    steps with a swap after each, ++step, writers (STEERER/WRITER order and
//    periods of :76-93, :144-180);
//  * run(): re-initialise, STEERER_INITIALIZED, WRITER_INITIALIZED, steps
//    until maxSteps() or a steerer ends the simulation, STEERER_ALL_DONE
//    (:98-117);
//  * the update of every cell in z, y, x order through a neighbourhood that
//    resolves hood[Coord<DIM>] and hood[FixedCoord<X,Y,Z>()] relative to the
//    cell with the topology's rules (out of bounds -> edge cell; wrapped axes
//    folded), as VanillaUpdateFunctor / FixedNeighborhood do
//    (storage/vanillaupdatefunctor.h:12-35, storage/fixedneighborhood.h).
#ifndef LGD_ORACLE_SERIALSIMULATOR_H
#define LGD_ORACLE_SERIALSIMULATOR_H

#include <libgeodecomp/geometry/fixedcoord.h>
#include <libgeodecomp/parallelization/monolithicsimulator.h>
#include <libgeodecomp/storage/grid.h>

#include <utility>

namespace LibGeoDecomp {

template<typename CELL, typename GRID>
class OracleNeighborhood
{
public:
    static const int DIM = GRID::DIM;

    OracleNeighborhood(const GRID& grid, const Coord<DIM>& center) : grid(grid), center(center) {}

    const CELL& operator[](const Coord<DIM>& rel) const { return grid[center + rel]; }

    template<int X, int Y, int Z>
    const CELL& operator[](FixedCoord<X, Y, Z>) const
    {
        Coord<DIM> c = center;
        c[0] += X;
        if (DIM > 1) {
            c[1 % DIM] += Y;
        }
        if (DIM > 2) {
            c[2 % DIM] += Z;
        }
        return grid[c];
    }

private:
    const GRID& grid;
    Coord<DIM> center;
};

template<typename CELL_TYPE>
class SerialSimulator : public MonolithicSimulator<CELL_TYPE>
{
public:
    typedef typename MonolithicSimulator<CELL_TYPE>::Topology Topology;
    typedef Grid<CELL_TYPE, Topology> GridType;
    typedef typename Steerer<CELL_TYPE>::SteererFeedback SteererFeedback;
    static const int DIM = Topology::DIM;
    static const unsigned NANO_STEPS = APITraits::SelectNanoSteps<CELL_TYPE>::VALUE;

    using MonolithicSimulator<CELL_TYPE>::chronometer;
    using MonolithicSimulator<CELL_TYPE>::initializer;
    using MonolithicSimulator<CELL_TYPE>::steerers;
    using MonolithicSimulator<CELL_TYPE>::stepNum;
    using MonolithicSimulator<CELL_TYPE>::writers;
    using MonolithicSimulator<CELL_TYPE>::gridDim;

    explicit SerialSimulator(Initializer<CELL_TYPE> *initializer) :
        MonolithicSimulator<CELL_TYPE>(initializer),
        curGrid(initializer->gridBox()),
        newGrid(initializer->gridBox())
    {
        stepNum = initializer->startStep();
        simArea << initializer->gridBox();
        initializer->grid(&curGrid);
        initializer->grid(&newGrid);
    }

    virtual void step()
    {
        SteererFeedback feedback;
        step(&feedback);
    }

    virtual void step(SteererFeedback *feedback)
    {
        TimeTotal t(&chronometer);
        handleInput(STEERER_NEXT_STEP, feedback);
        for (unsigned i = 0; i < NANO_STEPS; ++i) {
            nanoStep(i);
        }
        ++stepNum;
        handleOutput(stepNum == initializer->maxSteps() ? WRITER_ALL_DONE : WRITER_STEP_FINISHED);
    }

    virtual void run()
    {
        initializer->grid(&curGrid);
        stepNum = initializer->startStep();
        for (std::size_t i = 0; i < steerers.size(); ++i) {
            steerers[i]->setRegion(simArea);
        }
        SteererFeedback feedback;
        handleInput(STEERER_INITIALIZED, &feedback);
        handleOutput(WRITER_INITIALIZED);
        while (stepNum < initializer->maxSteps()) {
            if (feedback.simulationEnded()) {
                break;
            }
            step(&feedback);
        }
        handleInput(STEERER_ALL_DONE, &feedback);
    }

    virtual const typename Simulator<CELL_TYPE>::GridType *getGrid() { return &curGrid; }

    void nanoStep(unsigned nano)
    {
        TimeCompute t(&chronometer);
        // the edge cell of the target grid follows the source grid's
        newGrid.setEdge(curGrid.getEdge());
        const CoordBox<DIM> box = curGrid.boundingBox();
        for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
            OracleNeighborhood<CELL_TYPE, GridType> hood(curGrid, *i);
            CELL_TYPE& target = newGrid[*i];
            target.update(hood, nano);
        }
        std::swap(curGrid, newGrid);
    }

protected:
    GridType curGrid;
    GridType newGrid;
    Region<DIM> simArea;

    void handleOutput(WriterEvent event)
    {
        TimeOutput t(&chronometer);
        for (std::size_t i = 0; i < writers.size(); ++i) {
            if ((event != WRITER_STEP_FINISHED) || ((stepNum % writers[i]->getPeriod()) == 0)) {
                writers[i]->stepFinished(curGrid, stepNum, event);
            }
        }
    }

    void handleInput(SteererEvent event, SteererFeedback *feedback)
    {
        TimeInput t(&chronometer);
        for (std::size_t i = 0; i < steerers.size(); ++i) {
            if ((event != STEERER_NEXT_STEP) || (stepNum % steerers[i]->getPeriod() == 0)) {
                steerers[i]->nextStep(&curGrid, simArea, gridDim, stepNum, event, 0, true, feedback);
            }
        }
    }
};

}

#endif
