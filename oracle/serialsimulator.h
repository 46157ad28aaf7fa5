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
//  * the update path the reference's UpdateFunctor selector picks
//    (storage/updatefunctor.h:96-174):
//    - plain cells: update() of every cell in z, y, x order through a
//      neighbourhood that resolves hood[Coord<DIM>] and
//      hood[FixedCoord<X,Y,Z>()] relative to the cell with the topology's
//      rules (out of bounds -> edge cell; wrapped axes folded), as
//      VanillaUpdateFunctor does (storage/vanillaupdatefunctor.h:12-35);
//    - array-of-structs line cells (FixedCoordsOnly + UpdateLineX):
//      updateLineX(newLine, &x, endX, hood, nanoStep) per row, split into the
//      first cell, the middle and the last cell as
//      storage/linepointerupdatefunctor.h:121-148 splits it, hood[FixedCoord]
//      resolved at the shared x with the topology's rules;
//    - struct-of-arrays line cells (SoA + FixedCoordsOnly + UpdateLineX):
//      both grids are SoAGrids (edge radii on unwrapped axes only) and every
//      row runs updateLineX(hoodOld, indexEnd, hoodNew, nanoStep) through a
//      FixedNeighborhood with the wrap offsets of
//      storage/fixedneighborhoodupdatefunctor.h:126-258: Top/Bottom and
//      South/North set for the first/last row and plane of a wrapped axis,
//      the first and last cell of a wrapped x axis peeled off with
//      West/East set.  (One difference: before each of the peeled calls the
//      new grid's accessor is set to the old one's index, so a cell that
//      takes hoodNew by value still writes the right cells; the reference
//      carries that index only through a by-reference parameter.)  Like the
//      reference's, these offsets wrap only the first and last cell of an
//      axis, so wrapped axes need a stencil radius of 1.
#ifndef LGD_ORACLE_SERIALSIMULATOR_H
#define LGD_ORACLE_SERIALSIMULATOR_H

#include <libgeodecomp/geometry/fixedcoord.h>
#include <libgeodecomp/parallelization/monolithicsimulator.h>
#include <libgeodecomp/storage/fixedneighborhood.h>
#include <libgeodecomp/storage/grid.h>
#include <libgeodecomp/storage/soagrid.h>

#include <type_traits>
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

// hood of an array-of-structs line update: hood[FixedCoord<X,Y,Z>()] is the
// cell at (*x + X, y + Y, z + Z), x shared with the cell's loop
template<typename CELL, typename GRID>
class OracleLineNeighborhood
{
public:
    static const int DIM = GRID::DIM;

    OracleLineNeighborhood(const GRID& grid, const Coord<DIM>& rowOrigin, const long *x) :
        grid(grid), rowOrigin(rowOrigin), x(x)
    {}

    template<int X, int Y, int Z>
    const CELL& operator[](FixedCoord<X, Y, Z>) const
    {
        Coord<DIM> c = rowOrigin;
        c[0] += (int)*x + X;
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
    Coord<DIM> rowOrigin;
    const long *x;
};

template<typename CELL_TYPE>
class SerialSimulator : public MonolithicSimulator<CELL_TYPE>
{
public:
    typedef typename MonolithicSimulator<CELL_TYPE>::Topology Topology;
    static const bool SOA = APITraits::SelectSoAStreakUpdate<CELL_TYPE>::VALUE;
    static const bool LINE = APITraits::SelectLinePointerUpdate<CELL_TYPE>::VALUE;
    typedef typename std::conditional<SOA, SoAGrid<CELL_TYPE, Topology>, Grid<CELL_TYPE, Topology> >::type GridType;
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
        sweep(nano, std::integral_constant<int, SOA ? 2 : (LINE ? 1 : 0)>());
        std::swap(curGrid, newGrid);
    }

protected:
    GridType curGrid;
    GridType newGrid;
    Region<DIM> simArea;

    // update() cell by cell
    void sweep(unsigned nano, std::integral_constant<int, 0>)
    {
        const CoordBox<DIM> box = curGrid.boundingBox();
        for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
            OracleNeighborhood<CELL_TYPE, GridType> hood(curGrid, *i);
            CELL_TYPE& target = newGrid[*i];
            target.update(hood, nano);
        }
    }

    // array-of-structs line update, rows in z, y order
    void sweep(unsigned nano, std::integral_constant<int, 1>)
    {
        const CoordBox<DIM> box = curGrid.boundingBox();
        const long endX = box.dimensions.x();
        for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
            if ((*i).x() != box.origin.x()) {
                continue;
            }
            CELL_TYPE *line = &newGrid[*i];
            long x = 0;
            OracleLineNeighborhood<CELL_TYPE, GridType> hood(curGrid, *i, &x);
            if (endX == 1) {
                CELL_TYPE::updateLineX(line, &x, endX, hood, nano);
                continue;
            }
            CELL_TYPE::updateLineX(line, &x, 1L, hood, nano);
            CELL_TYPE::updateLineX(line, &x, endX - 1, hood, nano);
            CELL_TYPE::updateLineX(line, &x, endX, hood, nano);
        }
    }

    // struct-of-arrays streak update, rows in z, y order
    void sweep(unsigned nano, std::integral_constant<int, 2>)
    {
        typedef typename APITraits::SelectStencil<CELL_TYPE>::Value Stencil;
        static_assert(Stencil::RADIUS == 1 || (!Topology::template WrapsAxis<0>::VALUE &&
                                               !Topology::template WrapsAxis<1>::VALUE &&
                                               !Topology::template WrapsAxis<2>::VALUE),
                      "struct-of-arrays cells with a wrapped axis need a stencil radius of 1");
        const Coord<3> dim = SoAGridHelpers::expand(curGrid.boundingBox().dimensions, 1);
        const SoALayout& layout = curGrid.soaLayout();
        const bool wx = Topology::wrapsAxis(0), wy = DIM > 1 && Topology::wrapsAxis(1),
                   wz = DIM > 2 && Topology::wrapsAxis(2);
        for (long z = 0; z < dim.z(); ++z) {
            for (long y = 0; y < dim.y(); ++y) {
                const long top = (wy && y == 0) ? dim.y() : 0;
                const long bottom = (wy && y == dim.y() - 1) ? -dim.y() : 0;
                const long south = (wz && z == 0) ? dim.z() : 0;
                const long north = (wz && z == dim.z() - 1) ? -dim.z() : 0;
                typename SoAGrid<CELL_TYPE, Topology>::Accessor hoodOld = curGrid.accessor(layout.index(0, y, z));
                typename SoAGrid<CELL_TYPE, Topology>::Accessor hoodNew = newGrid.accessor(layout.index(0, y, z));
                const long indexEnd = hoodOld.index() + dim.x();
                if (wx) {
                    FixedNeighborhood<CELL_TYPE> left(hoodOld, dim.x(), dim.x() == 1 ? -dim.x() : 0, top, bottom,
                                                      south, north);
                    long end = hoodOld.index() + 1;
                    CELL_TYPE::updateLineX(left, end, hoodNew, nano);
                    hoodNew.index() = hoodOld.index();
                }
                FixedNeighborhood<CELL_TYPE> hood(hoodOld, 0, 0, top, bottom, south, north);
                if (wx) {
                    long end = indexEnd - 1;
                    CELL_TYPE::updateLineX(hood, end, hoodNew, nano);
                    hoodNew.index() = hoodOld.index();
                    FixedNeighborhood<CELL_TYPE> right(hoodOld, 0, -dim.x(), top, bottom, south, north);
                    end = indexEnd;
                    CELL_TYPE::updateLineX(right, end, hoodNew, nano);
                } else {
                    long end = indexEnd;
                    CELL_TYPE::updateLineX(hood, end, hoodNew, nano);
                }
            }
        }
    }

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
