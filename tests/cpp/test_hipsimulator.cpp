// GPU tests of HipSimulator: the same user cells and plugins run through
// HipSimulator (kernels via the lgdhip C-ABI) and through the oracle
// SerialSimulator; grids at every writer callback and the plugin event
// sequences must agree (bit-exact for Life and Jacobi, FP32 tolerance for
// the LBM cell).  Mirrors the cases of
// src/libgeodecomp/parallelization/test/unit/cudasimulatortest.h:69-209 and
// serialsimulatortest.h:47-232 that apply to a single-GPU Cube simulator.
#include "cells.h"
#include "harness.h"

#include <libgeodecomp/io/mpiioinitializer.h>
#include <libgeodecomp/io/mpiiowriter.h>
#include <libgeodecomp/io/serialbovwriter.h>
#include <libgeodecomp/io/tracingwriter.h>
#include <libgeodecomp/misc/patternoptimizer.h>
#include <libgeodecomp/parallelization/autotuningsimulator.h>
#include <libgeodecomp/parallelization/hipsimulator.h>
#include <libgeodecomp/parallelization/nesting/hipslabstepper.h>
#include <serialsimulator.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <memory>

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

template<typename CELL, typename INIT>
void runBoth(INIT *(*make)(), unsigned writerPeriod, std::vector<Event> *evHip, std::vector<Event> *evRef,
             Snapshots<CELL> *snapHip, Snapshots<CELL> *snapRef, unsigned steererPeriod = 0,
             typename RecordingSteerer<CELL>::Hook steer = 0)
{
    HipSimulator<CELL> hip(make());
    SerialSimulator<CELL> ref(make());
    hip.addWriter(new RecordingWriter<CELL>(evHip, writerPeriod, &Snapshots<CELL>::hook, snapHip));
    ref.addWriter(new RecordingWriter<CELL>(evRef, writerPeriod, &Snapshots<CELL>::hook, snapRef));
    if (steererPeriod) {
        hip.addSteerer(new RecordingSteerer<CELL>(evHip, steererPeriod, steer));
        ref.addSteerer(new RecordingSteerer<CELL>(evRef, steererPeriod, steer));
    }
    hip.run();
    ref.run();
    CHECK_EQ(hip.getStep(), ref.getStep());
}

// ---------------------------------------------------------------------
// Game of Life
// ---------------------------------------------------------------------
static ExampleLifeInitializer *makeExample() { return new ExampleLifeInitializer(800); }

LGD_TEST(life_reference_example_800_steps)
{
    std::vector<Event> eh, er;
    Snapshots<ConwayCell> sh, sr;
    runBoth<ConwayCell>(&makeExample, 10, &eh, &er, &sh, &sr);
    CHECK(eh == er);
    CHECK_EQ(sh.grids.size(), (std::size_t)(800 / 10 + 1));
    CHECK(sh.grids == sr.grids);
}

static RandomLifeInitializer *makeRandom() { return new RandomLifeInitializer(Coord<2>(517, 301), 95, 42, 0.3, 13); }

// plant a glider into the grid every time the steerer runs: the edits must
// reach the device (the reference CUDASimulator drops them)
static void plantGlider(GridBase<ConwayCell, 2> *grid, unsigned step, Steerer<ConwayCell>::SteererFeedback *,
                        void *)
{
    const int x0 = 20 + (int)(step * 7) % 400, y0 = 10 + (int)(step * 3) % 250;
    const int cells[5][2] = {{1, 0}, {2, 1}, {0, 2}, {1, 2}, {2, 2}};
    for (int y = 0; y < 5; ++y) {
        for (int x = 0; x < 5; ++x) {
            grid->set(Coord<2>(x0 + x - 1, y0 + y - 1), ConwayCell(false));
        }
    }
    for (const auto& c : cells) {
        grid->set(Coord<2>(x0 + c[0], y0 + c[1]), ConwayCell(true));
    }
}

LGD_TEST(life_random_with_steering)
{
    std::vector<Event> eh, er;
    Snapshots<ConwayCell> sh, sr;
    runBoth<ConwayCell>(&makeRandom, 3, &eh, &er, &sh, &sr, 5, &plantGlider);
    CHECK(eh == er);
    CHECK(!sh.grids.empty());
    CHECK(sh.grids == sr.grids);
}

static RandomLifeInitializer *makeLiveEdge()
{
    return new RandomLifeInitializer(Coord<2>(97, 1000), 70, 7, 0.3, 0, true);
}

LGD_TEST(life_alive_edge)
{
    std::vector<Event> eh, er;
    Snapshots<ConwayCell> sh, sr;
    runBoth<ConwayCell>(&makeLiveEdge, 35, &eh, &er, &sh, &sr);
    CHECK(eh == er);
    CHECK(sh.grids == sr.grids);
}

static void endAt19(GridBase<ConwayCell, 2> *, unsigned step, Steerer<ConwayCell>::SteererFeedback *fb, void *)
{
    if (step == 19) {
        fb->endSimulation();
    }
}

LGD_TEST(life_steerer_terminates)
{
    std::vector<Event> eh, er;
    Snapshots<ConwayCell> sh, sr;
    runBoth<ConwayCell>(&makeRandom, 4, &eh, &er, &sh, &sr, 1, &endAt19);
    CHECK(eh == er);
    CHECK(sh.grids == sr.grids);
}

// manual step() calls: the device state is downloaded lazily by getGrid()
LGD_TEST(life_manual_steps)
{
    HipSimulator<ConwayCell> hip(makeRandom());
    SerialSimulator<ConwayCell> ref(makeRandom());
    CHECK(*hip.getGrid() == *ref.getGrid());
    for (int i = 0; i < 9; ++i) {
        hip.step();
        ref.step();
        if (i % 4 == 0) {
            CHECK(*hip.getGrid() == *ref.getGrid());
        }
    }
    CHECK_EQ(hip.getStep(), 13u + 9u);
    CHECK(*hip.getGrid() == *ref.getGrid());
    CHECK_EQ(hip.sweepsIssued(), 9ull);
}

// run() twice re-initialises the grid (serialsimulatortest.h
// testRunMustResetGridPriorToSimulation)
LGD_TEST(life_run_twice)
{
    HipSimulator<ConwayCell> hip(makeExample());
    std::vector<Event> e1, e2;
    Snapshots<ConwayCell> s1, s2;
    hip.addWriter(new RecordingWriter<ConwayCell>(&e1, 100, &Snapshots<ConwayCell>::hook, &s1));
    hip.run();
    hip.addWriter(new RecordingWriter<ConwayCell>(&e2, 100, &Snapshots<ConwayCell>::hook, &s2));
    e1.clear();
    s1.grids.clear();
    hip.run();
    CHECK(e1 == e2);
    CHECK(s1.grids == s2.grids);
}

// a whole run without plugins is one batch of sweeps
LGD_TEST(life_batched_run)
{
    HipSimulator<ConwayCell> hip(makeExample());
    SerialSimulator<ConwayCell> ref(makeExample());
    hip.run();
    ref.run();
    CHECK_EQ(hip.sweepsIssued(), 800ull);
    CHECK(*hip.getGrid() == *ref.getGrid());
}

// tuned launch parameters do not change results
LGD_TEST(life_tuned_variants)
{
    SerialSimulator<ConwayCell> ref(makeRandom());
    ref.run();
    for (int variant = 0; variant <= 8; ++variant) {
        for (int spl : {1, 5, 32}) {
            if ((variant == 3 || variant == 4) && 2 * spl >= 64) {
                continue;
            }
            HipSimulator<ConwayCell> hip(makeRandom(), Coord<3>(), spl, variant);
            hip.run();
            CHECK(*hip.getGrid() == *ref.getGrid());
        }
    }
}

// ---------------------------------------------------------------------
// Jacobi3D FP64
// ---------------------------------------------------------------------
static JacobiInitializer *makeJacobi() { return new JacobiInitializer(Coord<3>(64, 48, 40), 30, 3, 0.25); }

LGD_TEST(jacobi_bit_exact)
{
    std::vector<Event> eh, er;
    Snapshots<JacobiCell> sh, sr;
    runBoth<JacobiCell>(&makeJacobi, 7, &eh, &er, &sh, &sr);
    CHECK(eh == er);
    CHECK_EQ(sh.grids.size(), (std::size_t)(1 + 4 + 1));
    CHECK(sh.grids == sr.grids);
}

// ---------------------------------------------------------------------
// periodic topologies (Torus, and axes wrapped selectively): Jacobi runs
// its blocked sweeps with wrapped tile coordinates (leftover steps: single
// sweeps after a periodic halo refresh), Life its bit-packed blocked
// generations with wrapped rows and word columns (a wrapped x axis of
// nx % 32 != 0: byte-grid generations after a halo refresh); identical to
// the oracle at every writer callback
// ---------------------------------------------------------------------
template<typename CELL> CELL makeLife(double u) { return CELL(u < 0.3); }
template<typename CELL> CELL makeTemp(double u) { return CELL(u); }

static RandomCellInitializer<TorusLifeCell, 2> *makeTorusLife()
{
    return new RandomCellInitializer<TorusLifeCell, 2>(Coord<2>(131, 77), 40, 11, &makeLife<TorusLifeCell>,
                                                       TorusLifeCell(true));
}
static RandomCellInitializer<CylinderLifeCell, 2> *makeCylinderLife()
{
    return new RandomCellInitializer<CylinderLifeCell, 2>(Coord<2>(96, 45), 33, 12, &makeLife<CylinderLifeCell>,
                                                          CylinderLifeCell(true));
}
static RandomCellInitializer<TorusJacobiCell, 3> *makeTorusJacobi()
{
    return new RandomCellInitializer<TorusJacobiCell, 3>(Coord<3>(37, 22, 15), 12, 13, &makeTemp<TorusJacobiCell>,
                                                         TorusJacobiCell(0.5));
}
static RandomCellInitializer<SlabJacobiCell, 3> *makeSlabJacobi()
{
    return new RandomCellInitializer<SlabJacobiCell, 3>(Coord<3>(30, 17, 21), 9, 14, &makeTemp<SlabJacobiCell>,
                                                        SlabJacobiCell(0.25));
}

LGD_TEST(life_torus)
{
    std::vector<Event> eh, er;
    Snapshots<TorusLifeCell> sh, sr;
    runBoth<TorusLifeCell>(&makeTorusLife, 6, &eh, &er, &sh, &sr);
    CHECK(eh == er);
    CHECK(sh.grids == sr.grids);
}

// a torus whose rows are whole packed words: the blocked kernels load the
// periodic images (a writer period of 33: launches of 31 and 2 generations)
static RandomCellInitializer<TorusLifeCell, 2> *makeTorusLifePacked()
{
    return new RandomCellInitializer<TorusLifeCell, 2>(Coord<2>(128, 77), 70, 15, &makeLife<TorusLifeCell>,
                                                       TorusLifeCell(true));
}

LGD_TEST(life_torus_packed)
{
    std::vector<Event> eh, er;
    Snapshots<TorusLifeCell> sh, sr;
    runBoth<TorusLifeCell>(&makeTorusLifePacked, 33, &eh, &er, &sh, &sr);
    CHECK(eh == er);
    CHECK(sh.grids == sr.grids);
}

LGD_TEST(life_wrapped_x_only)
{
    std::vector<Event> eh, er;
    Snapshots<CylinderLifeCell> sh, sr;
    runBoth<CylinderLifeCell>(&makeCylinderLife, 5, &eh, &er, &sh, &sr);
    CHECK(sh.grids == sr.grids);
}

LGD_TEST(jacobi_torus)
{
    std::vector<Event> eh, er;
    Snapshots<TorusJacobiCell> sh, sr;
    runBoth<TorusJacobiCell>(&makeTorusJacobi, 4, &eh, &er, &sh, &sr);
    CHECK(eh == er);
    CHECK(sh.grids == sr.grids);
}

LGD_TEST(jacobi_wrapped_x_and_z)
{
    std::vector<Event> eh, er;
    Snapshots<SlabJacobiCell> sh, sr;
    runBoth<SlabJacobiCell>(&makeSlabJacobi, 3, &eh, &er, &sh, &sr);
    CHECK(sh.grids == sr.grids);
}

// a writer period of 9: each run holds two 4-sweep blocked launches with
// wrapped tile coordinates and one halo-refresh single sweep
LGD_TEST(jacobi_wrapped_x_and_z_blocked)
{
    std::vector<Event> eh, er;
    Snapshots<SlabJacobiCell> sh, sr;
    runBoth<SlabJacobiCell>(&makeSlabJacobi, 9, &eh, &er, &sh, &sr);
    CHECK(eh == er);
    CHECK(sh.grids == sr.grids);
}

// the BOV snapshots of a HipSimulator run equal the oracle simulator's byte
// for byte (headers up to the file-name prefix)
static std::string slurp(const std::string& name)
{
    std::ifstream in(name.c_str(), std::ios::binary);
    std::stringstream buf;
    buf << in.rdbuf();
    return buf.str();
}

LGD_TEST(jacobi_bov_snapshots)
{
    {
        HipSimulator<JacobiCell> hip(makeJacobi());
        hip.addWriter(new SerialBOVWriter<JacobiCell>(Selector<JacobiCell>(&JacobiCell::temp, "temp"), "hipbov", 7));
        hip.run();
        SerialSimulator<JacobiCell> ref(makeJacobi());
        ref.addWriter(new SerialBOVWriter<JacobiCell>(Selector<JacobiCell>(&JacobiCell::temp, "temp"), "refbov", 7));
        ref.run();
    }
    for (unsigned step : {0u, 7u, 14u, 21u, 28u, 30u}) {
        char a[64], b[64];
        std::snprintf(a, sizeof(a), "hipbov.%05u", step);
        std::snprintf(b, sizeof(b), "refbov.%05u", step);
        const std::string da = slurp(std::string(a) + ".data"), db = slurp(std::string(b) + ".data");
        CHECK(!da.empty());
        CHECK(da == db);
        std::string ha = slurp(std::string(a) + ".bov"), hb = slurp(std::string(b) + ".bov");
        CHECK(!ha.empty());
        ha.replace(ha.find("hipbov"), 6, "refbov");
        CHECK(ha == hb);
        for (const char *stem : {a, b}) {
            std::remove((std::string(stem) + ".bov").c_str());
            std::remove((std::string(stem) + ".data").c_str());
        }
    }
}

static void heatCorner(GridBase<JacobiCell, 3> *grid, unsigned, Steerer<JacobiCell>::SteererFeedback *, void *)
{
    grid->set(Coord<3>(1, 2, 3), JacobiCell(100.0));
}

LGD_TEST(jacobi_steered)
{
    std::vector<Event> eh, er;
    Snapshots<JacobiCell> sh, sr;
    runBoth<JacobiCell>(&makeJacobi, 10, &eh, &er, &sh, &sr, 6, &heatCorner);
    CHECK(eh == er);
    CHECK(sh.grids == sr.grids);
}

// ---------------------------------------------------------------------
// D3Q19 LBM, FP32 device vs FP32 and FP64 oracle cells
// ---------------------------------------------------------------------
LGD_TEST(lbm_cavity)
{
    const Coord<3> dim(24, 20, 16);
    const unsigned steps = 40;
    HipSimulator<LBMCavityCell<float> > hip(new CavityInitializer<float>(dim, steps, 9));
    SerialSimulator<LBMCavityCell<float> > ref32(new CavityInitializer<float>(dim, steps, 9));
    SerialSimulator<LBMCavityCell<double> > ref64(new CavityInitializer<double>(dim, steps, 9));
    hip.run();
    ref32.run();
    ref64.run();
    double err32 = 0, err64 = 0, errRho = 0, errJ = 0, errRho32 = 0;
    double mass0 = 0, mass = 0;
    CavityInitializer<double> init(dim, steps, 9);
    Grid<LBMCavityCell<double>, Topologies::Cube<3>::Topology> g0(CoordBox<3>(Coord<3>(), dim));
    init.grid(&g0);
    CoordBox<3> box(Coord<3>(), dim);
    typedef LBMCavityCell<double> C;
    for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
        const LBMCavityCell<float> a = hip.getGrid()->get(*i);
        const LBMCavityCell<float> b = ref32.getGrid()->get(*i);
        const LBMCavityCell<double> c = ref64.getGrid()->get(*i);
        double ra = 0, rc = 0, rb = 0, ja[3] = {0, 0, 0}, jc[3] = {0, 0, 0};
        for (int q = 0; q < 19; ++q) {
            err32 = std::max(err32, (double)std::fabs(a.f[q] - b.f[q]));
            err64 = std::max(err64, std::fabs((double)a.f[q] - c.f[q]));
            ra += a.f[q];
            rb += b.f[q];
            rc += c.f[q];
            mass += a.f[q];
            mass0 += g0[*i].f[q];
            ja[0] += C::EX[q] * (double)a.f[q];
            ja[1] += C::EY[q] * (double)a.f[q];
            ja[2] += C::EZ[q] * (double)a.f[q];
            jc[0] += C::EX[q] * c.f[q];
            jc[1] += C::EY[q] * c.f[q];
            jc[2] += C::EZ[q] * c.f[q];
        }
        errRho = std::max(errRho, std::fabs(ra - rc));
        errRho32 = std::max(errRho32, std::fabs(rb - rc));
        for (int k = 0; k < 3; ++k) {
            errJ = std::max(errJ, std::fabs(ja[k] - jc[k]));
        }
    }
    std::printf("  lbm %dx%dx%d x %u: max|f_gpu - f_oracle64| %.3e, max|f_gpu - f_oracle32| %.3e, "
                "max|rho - rho64| %.3e (FP32 cell on the CPU: %.3e), max|j - j64| %.3e, mass drift %.3e\n",
                dim.x(), dim.y(), dim.z(), steps, err64, err32, errRho, errRho32, errJ, std::fabs(mass - mass0));
    // the device keeps shifted FP32 populations (g = f - w): more accurate
    // than the same cell in plain FP32 on the CPU, which drifts by about one
    // ulp of rho per step; the values handed back are rounded to float f
    CHECK(err64 < 2e-7);
    CHECK(errRho < 2e-6 && errRho <= errRho32);
    CHECK(errJ < 1e-6);
    CHECK(std::fabs(mass - mass0) < 1e-7 * mass0 + 19 * dim.prod() * std::ldexp(1.0, -25));
}

// ---------------------------------------------------------------------
// restart from an MPI-IO snapshot: the resumed HipSimulator run ends with
// the uninterrupted run's grid (and the oracle's)
// ---------------------------------------------------------------------
template<typename CELL>
void restartCase(std::function<Initializer<CELL> *()> make, unsigned at, const std::string& prefix,
                 std::function<bool(const CELL&, const CELL&)> same)
{
    typedef typename HipSimulator<CELL>::Topology Topology;
    static const int DIM = Topology::DIM;
    Initializer<CELL> *first = make();
    const unsigned total = first->maxSteps();
    HipSimulator<CELL> full(first);
    full.addWriter(new MPIIOWriter<CELL>(prefix, at, total));
    full.run();
    char name[256];
    std::snprintf(name, sizeof(name), "%s%05u.mpiio", prefix.c_str(), at);
    HipSimulator<CELL> resumed(new MPIIOInitializer<CELL>(name));
    resumed.run();
    SerialSimulator<CELL> ref(make());
    ref.run();
    CHECK_EQ(resumed.getStep(), total);
    int differFull = 0, differRef = 0;
    CoordBox<DIM> box = full.getGrid()->boundingBox();
    for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
        const CELL a = full.getGrid()->get(*i), b = resumed.getGrid()->get(*i), c = ref.getGrid()->get(*i);
        differFull += !same(a, b);
        differRef += !same(b, c);
    }
    CHECK_EQ(differFull, 0);
    CHECK_EQ(differRef, 0);
}

LGD_TEST(restart_from_snapshot_families)
{
    restartCase<ConwayCell>([]() { return new RandomLifeInitializer(Coord<2>(517, 301), 90, 42, 0.3); }, 40,
                            "life", [](const ConwayCell& a, const ConwayCell& b) { return a == b; });
    restartCase<TorusLifeCell>(
        []() {
            return new RandomCellInitializer<TorusLifeCell, 2>(
                Coord<2>(131, 77), 50, 3, [](double u) { return TorusLifeCell(u < 0.3); }, TorusLifeCell(false));
        },
        20, "torus", [](const TorusLifeCell& a, const TorusLifeCell& b) { return a.alive == b.alive; });
    restartCase<JacobiCell>([]() { return new JacobiInitializer(Coord<3>(40, 33, 29), 26, 9, 0.5); }, 13,
                            "jacobi", [](const JacobiCell& a, const JacobiCell& b) { return a == b; });
    // LBM: the host cell holds f, the device g = f - w; a snapshot of f
    // re-uploaded rounds once more, so agreement is to FP32 rounding
    restartCase<LBMCavityCell<float> >(
        []() { return new CavityInitializer<float>(Coord<3>(24, 20, 16), 30, 9); }, 10, "lbm",
        [](const LBMCavityCell<float>& a, const LBMCavityCell<float>& b) {
            for (int q = 0; q < 19; ++q) {
                if (std::fabs(a.f[q] - b.f[q]) > 2e-6f) {
                    return false;
                }
            }
            return a.flags == b.flags;
        });
}

// ---------------------------------------------------------------------
// the autotuner over HipSimulator (misc/cudasimulationfactory.h and
// parallelization/autotuningsimulator.h counterparts)
// ---------------------------------------------------------------------
template<typename CELL, int DIM>
bool sameGrid(const GridBase<CELL, DIM> *a, const GridBase<CELL, DIM> *b)
{
    CoordBox<DIM> box = a->boundingBox();
    for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
        if (!(a->get(*i) == b->get(*i))) {
            return false;
        }
    }
    return true;
}

// ---------------------------------------------------------------------
// slab stepper: z slabs with ghost zones, one blocked launch per slab and
// round split into rim and inner planes, rims copied to the neighbours'
// ghost zones on a copy stream (the multi-GPU step loop on one GPU)
// ---------------------------------------------------------------------
static void slabCase(Coord<3> dim, int slabs, int width, unsigned steps, double edgeValue)
{
    std::printf("  slabs %d, ghost zone %d, %u steps, %dx%dx%d\n", slabs, width, steps, dim.x(), dim.y(), dim.z());
    typedef HipSlabStepper<JacobiCell>::GridType JacobiGrid;
    JacobiInitializer init(dim, steps, 7, edgeValue);
    JacobiGrid initial(CoordBox<3>(Coord<3>(), dim));
    init.grid(&initial);
    HipSlabStepper<JacobiCell> stepper(initial, slabs, width);
    CHECK_EQ(stepper.slabs(), (std::size_t)slabs);
    stepper.update(steps);
    JacobiGrid got = initial;
    stepper.grid(&got);
    SerialSimulator<JacobiCell> ref(new JacobiInitializer(dim, steps, 7, edgeValue));
    ref.run();
    CHECK(sameGrid(&got, ref.getGrid()));
    CHECK_EQ(stepper.currentStep(), steps);
    const unsigned long long rounds = (steps + width - 1) / width;
    CHECK_EQ(stepper.roundsRun(), rounds);
    CHECK_EQ(stepper.planesExchanged(), rounds * 2 * (unsigned long long)(slabs - 1) * width);
    // every slab updates its rim next to each neighbour separately from its
    // inner planes unless the slab is too thin to have inner planes
    CHECK(stepper.rimUpdates() <= rounds * 2 * (unsigned long long)(slabs - 1));
    // the stepper continues from its state: more steps equal a longer run
    stepper.update(3);
    stepper.grid(&got);
    SerialSimulator<JacobiCell> ref2(new JacobiInitializer(dim, steps + 3, 7, edgeValue));
    ref2.run();
    CHECK(sameGrid(&got, ref2.getGrid()));
}

LGD_TEST(slab_stepper_equals_oracle)
{
    slabCase(Coord<3>(45, 38, 40), 1, 4, 9, 0.25);  // one slab: no exchange
    slabCase(Coord<3>(45, 38, 40), 2, 4, 12, 0.25);
    slabCase(Coord<3>(64, 20, 37), 3, 4, 11, 0.5);  // uneven slabs, a partial round
    slabCase(Coord<3>(30, 33, 16), 4, 4, 8, 0.0);   // slabs one ghost zone thick: no inner planes
    slabCase(Coord<3>(30, 33, 21), 5, 2, 7, 0.5);   // two-sweep rounds
    slabCase(Coord<3>(17, 9, 12), 4, 3, 10, 0.25);  // three-sweep rounds
    slabCase(Coord<3>(17, 9, 12), 6, 1, 5, 0.25);   // one sweep per round (the one-sweep tile kernel)
    slabCase(Coord<3>(130, 70, 96), 3, 4, 8, 0.0);  // the two-rows-per-thread default for larger slabs
}

// the stepper's rate against the single-grid blocked sweeps on the same
// 256^3 state (printed, not asserted: a measurement, not a contract)
LGD_TEST(slab_stepper_rate)
{
    typedef HipSlabStepper<JacobiCell>::GridType JacobiGrid;
    const Coord<3> dim(256, 256, 256);
    const unsigned steps = 40;
    JacobiInitializer init(dim, steps, 9, 0.0);
    JacobiGrid initial(CoordBox<3>(Coord<3>(), dim));
    init.grid(&initial);
    const double cells = (double)dim.x() * dim.y() * dim.z();
    // single grid: lgdhip_jacobi3d_f64_run on two device buffers
    lgdhip_grid g;
    CHECK_EQ(lgdhip_grid_make(&g, 3, dim.x(), dim.y(), dim.z(), 1, 8, 1), 0);
    HipKernels::DeviceMemory a((std::size_t)lgdhip_grid_bytes(&g)), b((std::size_t)lgdhip_grid_bytes(&g));
    const double zero = 0.0;
    CHECK_EQ(lgdhip_grid_fill_edge(a.get(), &g, &zero, nullptr), 0);
    CHECK_EQ(lgdhip_grid_fill_edge(b.get(), &g, &zero, nullptr), 0);
    int inB = 0;
    CHECK_EQ(lgdhip_jacobi3d_f64_run((double *)a.get(), (double *)b.get(), &g, steps, nullptr, &inB, nullptr), 0);
    CHECK_EQ(lgdhip_device_synchronize(), 0);
    auto t0 = std::chrono::steady_clock::now();
    CHECK_EQ(lgdhip_jacobi3d_f64_run((double *)a.get(), (double *)b.get(), &g, steps, nullptr, &inB, nullptr), 0);
    CHECK_EQ(lgdhip_device_synchronize(), 0);
    const double single = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int slabs : {2, 4, 8}) {
        HipSlabStepper<JacobiCell> stepper(initial, slabs, 4);
        stepper.update(steps);  // warm-up
        stepper.sync();
        auto t1 = std::chrono::steady_clock::now();
        stepper.update(steps);
        stepper.sync();
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        std::printf("  256^3 x %u steps: single grid %.1f GLUPS, %d slabs %.1f GLUPS (%llu planes exchanged)\n", steps,
                    cells * steps / single * 1e-9, slabs, cells * steps / dt * 1e-9,
                    stepper.planesExchanged() / 2);
        CHECK(dt > 0);
    }
}

LGD_TEST(slab_stepper_rejects_thin_slabs)
{
    HipSlabStepper<JacobiCell>::GridType initial(CoordBox<3>(Coord<3>(), Coord<3>(8, 8, 10)));
    bool thrown = false;
    try {
        HipSlabStepper<JacobiCell> stepper(initial, 3, 4);  // slabs of 3 planes, ghost zones of 4
    } catch (const std::invalid_argument&) {
        thrown = true;
    }
    CHECK(thrown);
}

LGD_TEST(hip_simulation_factory_every_jacobi_variant)
{
    const Coord<3> dim(45, 38, 31);
    std::shared_ptr<VarStepInitializerProxy<JacobiCell> > proxy(
        new VarStepInitializerProxy<JacobiCell>(new JacobiInitializer(dim, 19, 4, 0.25)));
    HipSimulationFactory<JacobiCell> factory(proxy);
    CHECK(factory.parameters().contains("Variant"));
    SerialSimulator<JacobiCell> ref(proxy->clone());
    ref.run();
    SimulationParameters p = factory.parameters();
    for (int k = 0; k < 13; ++k) {  // every index of the Variant set
        p["Variant"].setValue(k);
        const double fit = factory(p);
        CHECK(fit < 0 && fit > -60);
        factory.parameters()["Variant"].setValue(k);
        std::unique_ptr<Simulator<JacobiCell> > sim(factory());
        sim->run();
        CHECK(sameGrid(dynamic_cast<MonolithicSimulator<JacobiCell> *>(sim.get())->getGrid(), ref.getGrid()));
    }
}

LGD_TEST(hip_simulation_factory_rejected_configuration_rates_worst)
{
    std::shared_ptr<VarStepInitializerProxy<ConwayCell> > proxy(
        new VarStepInitializerProxy<ConwayCell>(new RandomLifeInitializer(Coord<2>(300, 200), 40, 8, 0.3)));
    HipSimulationFactory<ConwayCell> factory(proxy);
    SimulationParameters p = factory.parameters();
    CHECK(factory(p) < 0 && factory(p) > -60);
    // 32 generations per launch need more than the 64 register rows of
    // variant 3: the family rejects it (before any launch) with a runtime
    // error, which rates the worst fitness
    p["StepsPerLaunch"].setValue(31);
    p["Variant"].setValue(2);
    CHECK_EQ((int)p["StepsPerLaunch"], 32);
    CHECK_EQ((int)p["Variant"], 3);
    CHECK_EQ(factory(p), std::numeric_limits<double>::lowest());
}

LGD_TEST(autotuning_simulator_jacobi_and_life)
{
    std::vector<Event> evTuned, evRef;
    Snapshots<JacobiCell> sTuned, sRef;
    AutoTuningSimulator<JacobiCell, PatternOptimizer> tuned(new JacobiInitializer(Coord<3>(64, 64, 64), 40, 2, 0.5),
                                                           6);
    tuned.addWriter(new RecordingWriter<JacobiCell>(&evTuned, 10, &Snapshots<JacobiCell>::hook, &sTuned));
    tuned.run();
    SerialSimulator<JacobiCell> ref(new JacobiInitializer(Coord<3>(64, 64, 64), 40, 2, 0.5));
    ref.addWriter(new RecordingWriter<JacobiCell>(&evRef, 10, &Snapshots<JacobiCell>::hook, &sRef));
    ref.run();
    std::printf("  Jacobi 64^3: best %s after tuning runs of %u steps (fitness %g)\n",
                tuned.bestParameters().toString().c_str(), tuned.stepsPerTuningRun(), tuned.bestTuningFitness());
    CHECK(evTuned == evRef);
    CHECK(sTuned.grids == sRef.grids);
    CHECK_EQ(tuned.getStep(), 40u);
    CHECK(tuned.stepsPerTuningRun() >= 1u && tuned.stepsPerTuningRun() <= 40u);

    AutoTuningSimulator<ConwayCell, PatternOptimizer> life(new RandomLifeInitializer(Coord<2>(1024, 768), 300, 6),
                                                           5);
    life.run();
    SerialSimulator<ConwayCell> lref(new RandomLifeInitializer(Coord<2>(1024, 768), 300, 6));
    lref.run();
    std::printf("  Life 1024x768: best %s\n", life.bestParameters().toString().c_str());
    CHECK(sameGrid(life.getGrid(), lref.getGrid()));
}

LGD_TEST(autotuning_cache_reuses_parameters)
{
    // the first run tunes and stores its result; a second run of the same
    // cell, grid and device takes it from the file without tuning and still
    // equals the oracle; another grid is another key
    const std::string path = "tuning_cache.txt";
    std::remove(path.c_str());
    AutoTuningSimulator<JacobiCell, PatternOptimizer> first(new JacobiInitializer(Coord<3>(48, 40, 32), 20, 2, 0.5),
                                                            4);
    first.setCacheFile(path);
    first.run();
    CHECK(!first.usedCache());
    CHECK(first.stepsPerTuningRun() >= 1u);
    AutoTuningSimulator<JacobiCell, PatternOptimizer> second(new JacobiInitializer(Coord<3>(48, 40, 32), 20, 2, 0.5),
                                                             4);
    second.setCacheFile(path);
    second.run();
    CHECK(second.usedCache());
    CHECK_EQ(second.stepsPerTuningRun(), 0u);
    CHECK(second.bestParameters().toString() == first.bestParameters().toString());
    std::printf("  cached %s\n", second.bestParameters().toString().c_str());
    SerialSimulator<JacobiCell> ref(new JacobiInitializer(Coord<3>(48, 40, 32), 20, 2, 0.5));
    ref.run();
    CHECK(sameGrid(second.getGrid(), ref.getGrid()));
    AutoTuningSimulator<JacobiCell, PatternOptimizer> other(new JacobiInitializer(Coord<3>(40, 40, 32), 20, 2, 0.5),
                                                            2);
    other.setCacheFile(path);
    other.run();
    CHECK(!other.usedCache());
    std::ifstream f(path.c_str());
    std::string line;
    int lines = 0;
    while (std::getline(f, line)) {
        ++lines;
        CHECK(line.find('\t') != std::string::npos);
    }
    CHECK_EQ(lines, 2);
    std::remove(path.c_str());
}

// ---------------------------------------------------------------------
// steerer edits travel as region uploads (only the changed streaks)
// ---------------------------------------------------------------------
template<typename CELL>
bool exactlyEqual(const CELL& a, const CELL& b)
{
    return a == b;
}

template<typename CELL, typename EDIT>
void regionUploadCase(Initializer<CELL> *(*make)(), unsigned steererPeriod, EDIT edit, bool expectRegion,
                      bool (*same)(const CELL&, const CELL&) = &exactlyEqual<CELL>)
{
    struct EditSteerer : public Steerer<CELL>
    {
        EDIT edit;
        EditSteerer(unsigned period, EDIT edit) : Steerer<CELL>(period), edit(edit) {}
        void nextStep(typename Steerer<CELL>::GridType *grid, const Region<Steerer<CELL>::Topology::DIM>&,
                      const Coord<Steerer<CELL>::Topology::DIM>&, unsigned step, SteererEvent event, std::size_t, bool,
                      typename Steerer<CELL>::SteererFeedback *)
        {
            if (event == STEERER_NEXT_STEP) {
                edit(grid, step);
            }
        }
    };
    HipSimulator<CELL> hip(make());
    SerialSimulator<CELL> ref(make());
    hip.addSteerer(new EditSteerer(steererPeriod, edit));
    ref.addSteerer(new EditSteerer(steererPeriod, edit));
    hip.run();
    ref.run();
    typedef typename HipSimulator<CELL>::Topology Topology;
    CoordBox<Topology::DIM> box = hip.getGrid()->boundingBox();
    int differ = 0;
    for (typename CoordBox<Topology::DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
        differ += !same(hip.getGrid()->get(*i), ref.getGrid()->get(*i));
    }
    std::printf("  full uploads %llu, region uploads %llu (%llu cells)\n", hip.fullUploads(), hip.regionUploads(),
                hip.regionCellsUploaded());
    CHECK_EQ(differ, 0);
    if (expectRegion) {
        CHECK_EQ(hip.fullUploads(), 1ull);
        CHECK(hip.regionUploads() > 0);
    } else {
        CHECK(hip.fullUploads() > 1);
    }
}

static Initializer<ConwayCell> *makeSteeredLife() { return new RandomLifeInitializer(Coord<2>(517, 301), 60, 42); }
static Initializer<TorusLifeCell> *makeSteeredTorus()
{
    return new RandomCellInitializer<TorusLifeCell, 2>(Coord<2>(131, 77), 30, 3,
                                                       [](double u) { return TorusLifeCell(u < 0.3); },
                                                       TorusLifeCell(false));
}
static Initializer<JacobiCell> *makeSteeredJacobi() { return new JacobiInitializer(Coord<3>(40, 33, 29), 24, 9, 0.5); }
static Initializer<LBMCavityCell<float> > *makeSteeredLBM() { return new CavityInitializer<float>(Coord<3>(24, 20, 16), 20, 9); }

LGD_TEST(steerer_region_uploads)
{
    // Life (bit-packed device state): a glider planted every 5 steps,
    // including streaks that cross 32-cell word boundaries
    regionUploadCase<ConwayCell>(&makeSteeredLife, 5, [](GridBase<ConwayCell, 2> *g, unsigned step) {
        plantGlider(g, step, 0, 0);
        Streak<2> s(Coord<2>(29, 100 + (int)step % 50), 70);
        std::vector<ConwayCell> row(41, ConwayCell(true));
        g->set(s, row.data());
    }, true);
    // Torus Life (byte device state), writes across the wrapped seam
    regionUploadCase<TorusLifeCell>(&makeSteeredTorus, 4, [](GridBase<TorusLifeCell, 2> *g, unsigned step) {
        for (int x = -2; x < 3; ++x) {
            g->set(Coord<2>(x, (int)step % 77), TorusLifeCell(true));
        }
    }, true);
    // Jacobi: a hot streak every 3 steps
    regionUploadCase<JacobiCell>(&makeSteeredJacobi, 3, [](GridBase<JacobiCell, 3> *g, unsigned step) {
        std::vector<JacobiCell> row(17, JacobiCell(2.0 + step));
        g->set(Streak<3>(Coord<3>(5, (int)step % 33, 7), 22), row.data());
    }, true);
    // LBM: populations and flags of one cell every 4 steps
    regionUploadCase<LBMCavityCell<float> >(&makeSteeredLBM, 4,
        [](GridBase<LBMCavityCell<float>, 3> *g, unsigned step) {
            const Coord<3> c(3 + (int)step % 15, 10, 8);
            LBMCavityCell<float> cell = g->get(c);
            cell.f[2] += 1e-3f;
            cell.f[3] -= 1e-3f;
            g->set(c, cell);
        }, true,
        [](const LBMCavityCell<float>& a, const LBMCavityCell<float>& b) {
            // FP32 populations against the FP32 oracle: within rounding
            for (int q = 0; q < 19; ++q) {
                if (std::fabs(a.f[q] - b.f[q]) > 1e-5f) {
                    return false;
                }
            }
            return a.flags == b.flags;
        });
    // an edge change needs a whole-grid upload
    regionUploadCase<JacobiCell>(&makeSteeredJacobi, 6, [](GridBase<JacobiCell, 3> *g, unsigned step) {
        g->setEdge(JacobiCell(0.1 * step));
    }, false);
}

// ---------------------------------------------------------------------
// writers read on demand: a writer that samples a few cells brings in only
// their x-rows (device -> host), a TracingWriter nothing at all
// ---------------------------------------------------------------------
template<typename CELL>
struct Samples
{
    typedef typename Writer<CELL>::GridType GridType;
    static const int DIM = GridType::DIM;
    std::vector<Coord<DIM> > at;
    std::vector<std::vector<CELL> > seen;

    static void hook(const GridType& g, unsigned, void *ctx)
    {
        Samples *s = static_cast<Samples *>(ctx);
        std::vector<CELL> cells;
        for (std::size_t i = 0; i < s->at.size(); ++i) {
            cells.push_back(g.get(s->at[i]));
        }
        s->seen.push_back(cells);
    }
};

// runs `make` on HipSimulator and the oracle with a sampling writer of
// period `period`; the samples must agree and the device must have sent only
// `rows` rows per callback after the first (whose grid is the host's own)
template<typename CELL, typename SAME>
void sampleCase(std::function<Initializer<CELL> *()> make, unsigned period,
                const std::vector<Coord<Samples<CELL>::DIM> >& at, std::size_t rowsPerCall, SAME same)
{
    std::vector<Event> eh, er;
    Samples<CELL> sh, sr;
    sh.at = sr.at = at;
    HipSimulator<CELL> hip(make());
    SerialSimulator<CELL> ref(make());
    hip.addWriter(new RecordingWriter<CELL>(&eh, period, &Samples<CELL>::hook, &sh));
    ref.addWriter(new RecordingWriter<CELL>(&er, period, &Samples<CELL>::hook, &sr));
    hip.run();
    ref.run();
    CHECK(eh == er);
    CHECK_EQ(sh.seen.size(), sr.seen.size());
    int differ = 0;
    for (std::size_t k = 0; k < sh.seen.size() && k < sr.seen.size(); ++k) {
        for (std::size_t i = 0; i < at.size(); ++i) {
            differ += !same(sh.seen[k][i], sr.seen[k][i]);
        }
    }
    CHECK_EQ(differ, 0);
    CHECK_EQ(hip.fullDownloads(), 0ull);
    CHECK_EQ(hip.rowsDownloaded(), rowsPerCall * (sh.seen.size() - 1));
    std::printf("  %zu writer calls, %zu rows downloaded, %llu whole-grid downloads\n", sh.seen.size(),
                hip.rowsDownloaded(), hip.fullDownloads());
}

LGD_TEST(writer_reads_only_its_rows)
{
    auto eq = [](const auto& a, const auto& b) { return a == b; };
    // Life, bit-packed device state: three cells in two rows (one of them
    // straddling a 32-cell word), an edge cell (no transfer)
    sampleCase<ConwayCell>([]() -> Initializer<ConwayCell> * { return makeRandom(); }, 10,
                           {Coord<2>(3, 4), Coord<2>(31, 4), Coord<2>(32, 4), Coord<2>(500, 290), Coord<2>(-1, 7)},
                           2, eq);
    // Torus Life, byte device state: a wrapped coordinate reads row ny - 1
    sampleCase<TorusLifeCell>([]() -> Initializer<TorusLifeCell> * { return makeTorusLife(); }, 6,
                              {Coord<2>(-1, -1), Coord<2>(130, 76), Coord<2>(5, 9)}, 2, eq);
    // Jacobi
    sampleCase<JacobiCell>([]() -> Initializer<JacobiCell> * { return makeJacobi(); }, 7,
                           {Coord<3>(0, 0, 0), Coord<3>(63, 47, 39), Coord<3>(10, 47, 39), Coord<3>(5, 5, 40)}, 2,
                           eq);
    // LBM: all 19 member planes of a row in one strided copy, FP32 rounding
    sampleCase<LBMCavityCell<float> >(
        []() -> Initializer<LBMCavityCell<float> > * { return makeSteeredLBM(); }, 5,
        {Coord<3>(1, 2, 3), Coord<3>(23, 19, 15), Coord<3>(0, 19, 15)}, 2,
        [](const LBMCavityCell<float>& a, const LBMCavityCell<float>& b) {
            for (int q = 0; q < 19; ++q) {
                if (std::fabs(a.f[q] - b.f[q]) > 1e-5f) {
                    return false;
                }
            }
            return true;
        });
}

LGD_TEST(writer_reading_everything_switches_to_one_download)
{
    // the snapshot writer reads every cell: the first rows come one by one,
    // the rest in one whole-grid transfer per callback
    std::vector<Event> eh;
    Snapshots<JacobiCell> sh;
    HipSimulator<JacobiCell> hip(makeJacobi());
    hip.addWriter(new RecordingWriter<JacobiCell>(&eh, 7, &Snapshots<JacobiCell>::hook, &sh));
    hip.run();
    const std::size_t calls = sh.grids.size() - 1;  // the first callback reads the host's own grid
    const std::size_t limit = std::max<std::size_t>(16, 48 * 40 / 32);
    CHECK_EQ(hip.fullDownloads(), (unsigned long long)calls);
    CHECK_EQ(hip.rowsDownloaded(), limit * calls);
}

LGD_TEST(tracing_writer_costs_no_transfer)
{
    std::ostringstream out;
    HipSimulator<ConwayCell> hip(makeRandom());
    hip.addWriter(new TracingWriter<ConwayCell>(5, 95, TracingWriter<ConwayCell>::OUTPUT_ON_ALL_RANKS, out));
    hip.run();
    const std::string log = out.str();
    std::size_t steps = 0;
    for (std::size_t p = log.find("TracingWriter::stepFinished()"); p != std::string::npos;
         p = log.find("TracingWriter::stepFinished()", p + 1)) {
        ++steps;
    }
    CHECK_EQ(steps, (std::size_t)16);   // the run starts at step 13: steps 15..90; 95 is ALL_DONE
    CHECK(log.find("TracingWriter::initialized()") != std::string::npos);
    CHECK(log.find("TracingWriter::allDone()") != std::string::npos);
    CHECK(log.find(" GLUPS") != std::string::npos);
    CHECK_EQ(hip.fullDownloads(), 0ull);
    CHECK_EQ(hip.rowsDownloaded(), (std::size_t)0);
}

int main(int argc, char **argv)
{
    int n = 0;
    if (lgdhip_get_device_count(&n) != 0 || n < 1) {
        std::printf("no HIP device\n");
        return 2;
    }
    return lgdtest::runAll(argc, argv);
}
