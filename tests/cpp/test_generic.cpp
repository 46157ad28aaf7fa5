// GPU tests of the generic kernel family: user cells without a hand-written
// family, their update() compiled by hipcc, run through HipSimulator and
// through the oracle SerialSimulator.  Grids at every writer callback and
// the plugin event sequences must agree exactly.  The ProbeCell cases mirror
// src/libgeodecomp/parallelization/test/unit/cudasimulatortest.h:69-209
// (1D 777, 2D 121 x 151, 3D 50 x 20 x 10; Cube and Torus; nano steps) with
// mixed topologies, radius-2 stencils and tiny tori added.
//
// Compiled with hipcc (tests/cpp/Makefile): the cells' update() runs on the
// GPU.
#include <algorithm>
#include "compare.h"
#include "devicecells.h"
#include "harness.h"

#include <libgeodecomp/hip/genericfamily.h>
#include <libgeodecomp/io/mpiioinitializer.h>
#include <libgeodecomp/io/mpiiowriter.h>
#include <libgeodecomp/misc/patternoptimizer.h>
#include <libgeodecomp/parallelization/autotuningsimulator.h>
#include <libgeodecomp/parallelization/hipsimulator.h>
#include <serialsimulator.h>

#include <chrono>
#include <cstdio>
#include <memory>

typedef Topologies::Cube<1>::Topology Line;
typedef Topologies::Torus<1>::Topology Ring;
typedef Topologies::Cube<2>::Topology Square;
typedef Topologies::Torus<2>::Topology Torus2;
typedef TopologiesHelpers::Topology<2, false, true, false> Cylinder2;
typedef Topologies::Cube<3>::Topology Cube3;
typedef Topologies::Torus<3>::Topology Torus3;
typedef TopologiesHelpers::Topology<3, true, false, true> Slab3;

// ---------------------------------------------------------------------
// the reference simulator test's shapes, Cube and Torus, nano steps
// ---------------------------------------------------------------------
LGD_TEST(probe_1d_cube_777)
{
    probeCase<ProbeCell<1, Line, 1, Stencils::Moore<1, 1> > >(Coord<1>(777), 13, 3);
}

LGD_TEST(probe_1d_torus_777_three_nano_steps)
{
    probeCase<ProbeCell<1, Ring, 3, Stencils::Moore<1, 1> > >(Coord<1>(777), 13, 4);
}

LGD_TEST(probe_2d_cube_121x151)
{
    probeCase<ProbeCell<2, Square, 1, Stencils::Moore<2, 1> > >(Coord<2>(121, 151), 11, 5);
}

LGD_TEST(probe_2d_torus_121x151_two_nano_steps)
{
    probeCase<ProbeCell<2, Torus2, 2, Stencils::Moore<2, 1> > >(Coord<2>(121, 151), 11, 5);
}

LGD_TEST(probe_2d_cylinder_von_neumann_radius_2)
{
    probeCase<ProbeCell<2, Cylinder2, 1, Stencils::VonNeumann<2, 2> > >(Coord<2>(67, 45), 9, 3);
}

LGD_TEST(probe_3d_cube_50x20x10)
{
    probeCase<ProbeCell<3, Cube3, 1, Stencils::Moore<3, 1> > >(Coord<3>(50, 20, 10), 7, 2);
}

LGD_TEST(probe_3d_torus_50x20x10_two_nano_steps)
{
    probeCase<ProbeCell<3, Torus3, 2, Stencils::Moore<3, 1> > >(Coord<3>(50, 20, 10), 7, 2);
}

LGD_TEST(probe_3d_slab_von_neumann_radius_2)
{
    probeCase<ProbeCell<3, Slab3, 1, Stencils::VonNeumann<3, 2> > >(Coord<3>(23, 17, 9), 5, 5);
}

// tori smaller than the stencil: the halo wraps around more than once
LGD_TEST(probe_tiny_tori)
{
    probeCase<ProbeCell<1, Ring, 1, Stencils::Moore<1, 3> > >(Coord<1>(2), 6, 2);
    probeCase<ProbeCell<2, Torus2, 1, Stencils::Moore<2, 2> > >(Coord<2>(3, 1), 6, 2);
    probeCase<ProbeCell<3, Torus3, 2, Stencils::Moore<3, 1> > >(Coord<3>(1, 2, 1), 4, 1);
}

// work-group shapes (HipSimulator's blockSize): same grids
LGD_TEST(probe_block_shapes)
{
    typedef ProbeCell<3, Torus3, 1, Stencils::Moore<3, 1> > Cell;
    probeCase<Cell>(Coord<3>(70, 9, 5), 4, 2, Coord<3>(128, 2, 1));
    probeCase<Cell>(Coord<3>(70, 9, 5), 4, 2, Coord<3>(64, 16, 1));
    probeCase<Cell>(Coord<3>(70, 9, 5), 4, 2, Coord<3>(256, 1, 1));
    typedef ProbeCell<1, Line, 1, Stencils::Moore<1, 2> > Cell1;
    probeCase<Cell1>(Coord<1>(1000), 5, 5, Coord<3>(64, 1, 1));
}

LGD_TEST(block_shape_rejected)
{
    typedef ProbeCell<2, Square, 1, Stencils::Moore<2, 1> > Cell;
    CHECK_THROWS(HipSimulator<Cell>(new ProbeInitializer<Cell>(Coord<2>(8, 8), 1, 1), Coord<3>(48, 1, 1)),
                 std::invalid_argument);
    CHECK_THROWS(HipSimulator<Cell>(new ProbeInitializer<Cell>(Coord<2>(8, 8), 1, 1), Coord<3>(512, 4, 1)),
                 std::invalid_argument);
}

// ---------------------------------------------------------------------
// steerer edits reach the device (the reference CUDASimulator drops them,
// cudasimulator.h:450-465)
// ---------------------------------------------------------------------
typedef ProbeCell<2, Torus2, 1, Stencils::Moore<2, 1> > SteeredCell;

static void nudge(GridBase<SteeredCell, 2> *grid, unsigned step, Steerer<SteeredCell>::SteererFeedback *, void *)
{
    const Coord<2> c((int)(step * 13) % 40, (int)(step * 7) % 30);
    SteeredCell cell = grid->get(c);
    cell.value += 1.0f;
    grid->set(c, cell);
}

LGD_TEST(probe_steerer_edits_uploaded)
{
    Snapshots<SteeredCell> sh, sr;
    const bool same = runBoth<SteeredCell>([]() { return new ProbeInitializer<SteeredCell>(Coord<2>(40, 30), 20, 7); },
                                           4, Coord<3>(), &sh, &sr, 3, &nudge);
    CHECK(same);
    checkProbes(sh, 20, 1);
}

// ---------------------------------------------------------------------
// the generic path against the hand-written families on the same update
// ---------------------------------------------------------------------
LGD_TEST(generic_life_equals_oracle_and_life2d_family)
{
    const Coord<2> dim(301, 217);
    const unsigned steps = 60;
    const uint64_t seed = 99;
    Snapshots<DeviceConwayCell> sh, sr;
    runBoth<DeviceConwayCell>(
        [&]() {
            return new RandomCellInitializer<DeviceConwayCell, 2>(
                dim, steps, seed, [](double u) { return DeviceConwayCell(u < 0.3); }, DeviceConwayCell(false));
        },
        10, Coord<3>(), &sh, &sr);
    // the registered Life2D family (bit-packed multi-generation kernels) on
    // the same start
    HipSimulator<ConwayCell> family(new RandomLifeInitializer(dim, steps, seed, 0.3));
    family.run();
    const GridBase<ConwayCell, 2> *g = family.getGrid();
    int differ = 0;
    std::size_t k = 0;
    CoordBox<2> box(Coord<2>(), dim);
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i, ++k) {
        differ += g->get(*i).alive != sh.grids.back()[k].alive;
    }
    CHECK_EQ(differ, 0);
}

LGD_TEST(generic_jacobi_equals_oracle_and_jacobi_family)
{
    const Coord<3> dim(37, 29, 23);
    const unsigned steps = 12;
    const uint64_t seed = 5;
    Snapshots<DeviceJacobiCell> sh, sr;
    runBoth<DeviceJacobiCell>(
        [&]() {
            return new RandomCellInitializer<DeviceJacobiCell, 3>(
                dim, steps, seed, [](double u) { return DeviceJacobiCell(u); }, DeviceJacobiCell(0.25));
        },
        4, Coord<3>(), &sh, &sr);
    HipSimulator<JacobiCell> family(new RandomCellInitializer<JacobiCell, 3>(
        dim, steps, seed, [](double u) { return JacobiCell(u); }, JacobiCell(0.25)));
    family.run();
    const GridBase<JacobiCell, 3> *g = family.getGrid();
    int differ = 0;
    std::size_t k = 0;
    CoordBox<3> box(Coord<3>(), dim);
    for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i, ++k) {
        differ += g->get(*i).temp != sh.grids.back()[k].temp;
    }
    CHECK_EQ(differ, 0);
}

// ---------------------------------------------------------------------
// updateHIP() / updateCUDA() replace update() on the GPU
// ---------------------------------------------------------------------
template<typename CELL>
void separateUpdateCase(int gpuPath)
{
    const Coord<2> dim(45, 31);
    const unsigned steps = 9;
    auto make = [&]() {
        return new RandomCellInitializer<CELL, 2>(
            dim, steps, 3, [](double u) { return CELL((int)(u * 1000)); }, CELL(0));
    };
    HipSimulator<CELL> hip(make());
    SerialSimulator<CELL> ref(make());
    hip.run();
    ref.run();
    int valueDiffers = 0, hipPath = 0, refPath = 0;
    CoordBox<2> box(Coord<2>(), dim);
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        const CELL a = hip.getGrid()->get(*i), b = ref.getGrid()->get(*i);
        valueDiffers += a.v != b.v;
        hipPath += a.path == gpuPath;
        refPath += b.path == 1;
    }
    CHECK_EQ(valueDiffers, 0);
    CHECK_EQ(hipPath, (int)box.size());
    CHECK_EQ(refPath, (int)box.size());
}

LGD_TEST(update_hip_runs_on_gpu)
{
    separateUpdateCase<HipUpdateCell>(2);
}

LGD_TEST(update_cuda_runs_on_gpu)
{
    separateUpdateCase<CudaUpdateCell>(3);
}

// ---------------------------------------------------------------------
// restart from an MPI-IO snapshot through the generic path: cycle counters
// and nano-step phase survive, grids end equal to the uninterrupted run's
// ---------------------------------------------------------------------
LGD_TEST(probe_restart_from_snapshot)
{
    typedef ProbeCell<3, Torus3, 2, Stencils::Moore<3, 1> > Cell;
    const Coord<3> dim(21, 13, 8);
    HipSimulator<Cell> full(new ProbeInitializer<Cell>(dim, 12, 4));
    full.addWriter(new MPIIOWriter<Cell>("probe", 5, 12));
    full.run();
    HipSimulator<Cell> resumed(new MPIIOInitializer<Cell>("probe00005.mpiio"));
    CHECK_EQ(resumed.getStep(), 5u);
    resumed.run();
    int differ = 0, invalid = 0;
    CoordBox<3> box(Coord<3>(), dim);
    for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
        const Cell a = full.getGrid()->get(*i), b = resumed.getGrid()->get(*i);
        differ += !(a == b);
        invalid += !b.valid || b.cycle != 24u;
    }
    CHECK_EQ(differ, 0);
    CHECK_EQ(invalid, 0);
}

// the autotuner over the generic family's work-group shapes
LGD_TEST(autotuning_generic_cells)
{
    typedef ProbeCell<2, Torus2, 2, Stencils::Moore<2, 1> > Cell;
    AutoTuningSimulator<Cell, PatternOptimizer> tuned(new ProbeInitializer<Cell>(Coord<2>(300, 200), 30, 8), 4);
    tuned.run();
    std::printf("  probe 300x200: best %s\n", tuned.bestParameters().toString().c_str());
    int invalid = 0;
    CoordBox<2> box(Coord<2>(), Coord<2>(300, 200));
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        const Cell c = tuned.getGrid()->get(*i);
        invalid += !c.valid || c.cycle != 60u;
    }
    CHECK_EQ(invalid, 0);
}

// ---------------------------------------------------------------------
// throughput of the generic path next to the hand-written families on the
// same update (printed, not asserted).  run() includes the initializer and
// the transfers, so the rate is taken from the difference of two runs of
// different lengths: cells x (s2 - s1) / (t2 - t1).
// ---------------------------------------------------------------------
template<typename CELL>
double runSeconds(Initializer<CELL> *init)
{
    HipSimulator<CELL> sim(init);
    sim.run();   // warm-up: module load, allocation, first upload
    const auto t0 = std::chrono::steady_clock::now();
    sim.run();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// GLUPS from the difference of two whole runs of s1 and s2 steps (the host
// initializer and the uploads cancel); the step spans below keep the GPU part
// of the difference well above that host time's jitter, and the best of three
// differences is taken
template<typename CELL>
double glups(std::function<Initializer<CELL> *(unsigned)> make, double cells, unsigned s1, unsigned s2)
{
    double best = 0.0;
    for (int rep = 0; rep < 3; ++rep) {
        const double t1 = runSeconds<CELL>(make(s1));
        const double t2 = runSeconds<CELL>(make(s2));
        if (t2 > t1) best = std::max(best, cells * (double)(s2 - s1) / (t2 - t1) * 1e-9);
    }
    return best;
}

LGD_TEST(generic_vs_family_throughput)
{
    const Coord<3> j(128, 128, 128);
    const double jc = 128.0 * 128 * 128;
    const double jg = glups<DeviceJacobiCell>(
        [&](unsigned s) {
            return new RandomCellInitializer<DeviceJacobiCell, 3>(
                j, s, 1, [](double u) { return DeviceJacobiCell(u); }, DeviceJacobiCell(0.0));
        },
        jc, 100, 8100);
    const double jf = glups<JacobiCell>(
        [&](unsigned s) {
            return new RandomCellInitializer<JacobiCell, 3>(
                j, s, 1, [](double u) { return JacobiCell(u); }, JacobiCell(0.0));
        },
        jc, 100, 8100);
    std::printf("  Jacobi3D FP64 128^3: generic %.1f GLUPS, Jacobi3DF64 family %.1f GLUPS\n", jg, jf);
    const Coord<2> l(4096, 4096);
    const double lc = 4096.0 * 4096;
    const double lg = glups<DeviceConwayCell>(
        [&](unsigned s) {
            return new RandomCellInitializer<DeviceConwayCell, 2>(
                l, s, 2, [](double u) { return DeviceConwayCell(u < 0.3); }, DeviceConwayCell(false));
        },
        lc, 200, 4200);
    const double lf = glups<ConwayCell>([&](unsigned s) { return new RandomLifeInitializer(l, s, 2, 0.3); }, lc, 200,
                                        20200);
    std::printf("  Life 4096^2: generic %.1f GLUPS, Life2D family %.1f GLUPS\n", lg, lf);
    CHECK(jg > 0 && jf > 0 && lg > 0 && lf > 0);
}

int main(int argc, char **argv)
{
    return lgdtest::runAll(argc, argv);
}
