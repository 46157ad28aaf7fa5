// CPU tests of the reference-compatible API layer and of the oracle
// SerialSimulator's plugin semantics, restating the expectations of
// src/libgeodecomp/parallelization/test/unit/serialsimulatortest.h.
#include "cells.h"
#include "devicecells.h"
#include "harness.h"
#include "soacells.h"

#include <libgeodecomp/geometry/region.h>
#include <libgeodecomp/io/mpiioinitializer.h>
#include <libgeodecomp/io/mpiiowriter.h>
#include <libgeodecomp/io/serialbovwriter.h>
#include <libgeodecomp/io/tracingwriter.h>
#include <libgeodecomp/io/varstepinitializerproxy.h>
#include <libgeodecomp/misc/patternoptimizer.h>
#include <libgeodecomp/misc/simulationfactory.h>
#include <libgeodecomp/storage/trackinggrid.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <serialsimulator.h>

// A cell that counts its updates and checks that every neighbour inside the
// grid has seen the same number of updates (after misc/testcell.h).
template<int DIM>
class CountingCell
{
public:
    class API : public APITraits::HasCubeTopology<DIM>
    {};

    explicit CountingCell(int cycle = 0, bool edge = false) : cycle(cycle), valid(true), edge(edge) {}

    template<typename HOOD>
    void update(const HOOD& hood, unsigned)
    {
        *this = hood[Coord<DIM>()];
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dx = -1; dx <= 1; ++dx) {
                Coord<DIM> c;
                c[0] = dx;
                if (DIM > 1) {
                    c[1 % DIM] = dy;
                }
                const CountingCell& n = hood[c];
                if (!n.edge && n.cycle != cycle) {
                    valid = false;
                }
            }
        }
        ++cycle;
    }

    bool operator==(const CountingCell& o) const { return cycle == o.cycle && valid == o.valid; }

    int cycle;
    bool valid;
    bool edge;
};

template<int DIM>
class CountingInitializer : public SimpleInitializer<CountingCell<DIM> >
{
public:
    CountingInitializer(Coord<DIM> dim, unsigned start, unsigned steps) :
        SimpleInitializer<CountingCell<DIM> >(dim, steps), start(start)
    {}

    virtual void grid(GridBase<CountingCell<DIM>, DIM> *ret)
    {
        ret->setEdge(CountingCell<DIM>(-1, true));
        CoordBox<DIM> box = ret->boundingBox();
        for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
            ret->set(*i, CountingCell<DIM>((int)start));
        }
    }

    unsigned startStep() const { return start; }

private:
    unsigned start;
};

LGD_TEST(geometry_basics)
{
    Coord<3> a(1, 2, 3), b(4, 5, 6);
    CHECK_EQ((a + b), Coord<3>(5, 7, 9));
    CHECK_EQ(Coord<3>(2, 3, 4).prod(), 24);
    CHECK_EQ(Coord<3>(1, 2, 3).toIndex(Coord<3>(10, 20, 30)), (std::size_t)(1 + 10 * 2 + 200 * 3));
    CoordBox<2> box(Coord<2>(3, 4), Coord<2>(5, 2));
    int n = 0;
    Coord<2> last;
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        CHECK(box.inBounds(*i));
        last = *i;
        ++n;
    }
    CHECK_EQ(n, 10);
    CHECK_EQ(last, Coord<2>(7, 5));
    CHECK_EQ((Stencils::VonNeumann<3, 1>::VOLUME), 7);
    CHECK_EQ((Stencils::VonNeumann<2, 2>::VOLUME), 13);
    CHECK_EQ((Stencils::Moore<2, 1>::VOLUME), 9);
    CHECK_EQ((Stencils::Moore<3, 1>::VOLUME), 27);
    typedef Topologies::Torus<2>::Topology Torus;
    typedef Topologies::Cube<2>::Topology Cube;
    CHECK_EQ(Torus::normalize(Coord<2>(-1, 12), Coord<2>(10, 12)), Coord<2>(9, 0));
    CHECK(Cube::isOutOfBounds(Coord<2>(-1, 3), Coord<2>(10, 12)));
    CHECK(!Torus::isOutOfBounds(Coord<2>(-1, 3), Coord<2>(10, 12)));
    CHECK((Torus::WrapsAxis<1>::VALUE) && !(Cube::WrapsAxis<0>::VALUE));
}

LGD_TEST(region_basics)
{
    Region<2> r;
    r << CoordBox<2>(Coord<2>(0, 0), Coord<2>(4, 3));
    r << Streak<2>(Coord<2>(2, 1), 7);
    CHECK_EQ(r.size(), (std::size_t)(12 + 3));
    CHECK_EQ(r.boundingBox(), CoordBox<2>(Coord<2>(0, 0), Coord<2>(7, 3)));
    CHECK(r.count(Coord<2>(6, 1)) && !r.count(Coord<2>(6, 0)));
}

LGD_TEST(apitraits_selection)
{
    CHECK((std::is_same<APITraits::SelectTopology<ConwayCell>::Value, Topologies::Cube<2>::Topology>::value));
    CHECK((std::is_same<APITraits::SelectStencil<ConwayCell>::Value, Stencils::Moore<2, 1> >::value));
    CHECK((std::is_same<APITraits::SelectTopology<JacobiCell>::Value, Topologies::Cube<3>::Topology>::value));
    CHECK((std::is_same<APITraits::SelectFixedCoordsOnlyUpdate<JacobiCell>::Value, APITraits::TrueType>::value));
    CHECK((std::is_same<APITraits::SelectFixedCoordsOnlyUpdate<ConwayCell>::Value, APITraits::FalseType>::value));
    CHECK((APITraits::SelectNanoSteps<ConwayCell>::VALUE == 1));
    CHECK((std::is_void<APITraits::SelectHipKernel<CountingCell<2> >::Value>::value));
}

LGD_TEST(writer_rejects_zero_period)
{
    std::vector<Event> ev;
    CHECK_THROWS(RecordingWriter<ConwayCell>(&ev, 0), std::invalid_argument);
}

// serialsimulatortest.h testSerialSimulatorShouldCallBackWriter: dim 17x12,
// start 13, end 21, writer period 3
LGD_TEST(serial_writer_events)
{
    std::vector<Event> ev;
    SerialSimulator<CountingCell<2> > sim(new CountingInitializer<2>(Coord<2>(17, 12), 13, 21));
    sim.addWriter(new RecordingWriter<CountingCell<2> >(&ev, 3));
    sim.run();
    std::vector<Event> exp;
    exp.push_back(Event{13, WRITER_INITIALIZED});
    for (unsigned i = 13 + 2; i < 21; i += 3) {
        exp.push_back(Event{i, WRITER_STEP_FINISHED});
    }
    exp.push_back(Event{21, WRITER_ALL_DONE});
    CHECK(ev == exp);
    CHECK_EQ(sim.getStep(), 21u);
}

// serialsimulatortest.h testSteererCallback: period 5
LGD_TEST(serial_steerer_events)
{
    std::vector<Event> ev;
    SerialSimulator<CountingCell<2> > sim(new CountingInitializer<2>(Coord<2>(17, 12), 13, 21));
    sim.addSteerer(new RecordingSteerer<CountingCell<2> >(&ev, 5));
    sim.run();
    std::vector<Event> exp;
    exp.push_back(Event{13, STEERER_INITIALIZED});
    for (unsigned t = 13; t < 21; ++t) {
        if (t % 5 == 0) {
            exp.push_back(Event{t, STEERER_NEXT_STEP});
        }
    }
    exp.push_back(Event{21, STEERER_ALL_DONE});
    CHECK(ev == exp);
}

// every cell sees neighbours of the same age at every step (testcell.h)
template<int DIM>
void countingRun(Coord<DIM> dim)
{
    SerialSimulator<CountingCell<DIM> > sim(new CountingInitializer<DIM>(dim, 0, 9));
    sim.run();
    const GridBase<CountingCell<DIM>, DIM> *g = sim.getGrid();
    CoordBox<DIM> box = g->boundingBox();
    for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
        CHECK(g->get(*i).valid);
        CHECK_EQ(g->get(*i).cycle, 9);
    }
}

LGD_TEST(serial_counting_cells)
{
    countingRun<2>(Coord<2>(17, 12));
    countingRun<3>(Coord<3>(7, 5, 4));
}

// testSteererCanTerminateSimulation: the run stops after the step during
// which a steerer requested the end
static void endAt19(GridBase<CountingCell<2>, 2> *, unsigned step, Steerer<CountingCell<2> >::SteererFeedback *fb,
                    void *)
{
    if (step == 19) {
        fb->endSimulation();
    }
}

LGD_TEST(serial_steerer_terminates)
{
    std::vector<Event> ev;
    SerialSimulator<CountingCell<2> > sim(new CountingInitializer<2>(Coord<2>(17, 12), 13, 30));
    sim.addSteerer(new RecordingSteerer<CountingCell<2> >(&ev, 1, endAt19));
    sim.run();
    CHECK_EQ(sim.getStep(), 20u);
    CHECK_EQ(sim.getGrid()->get(Coord<2>(3, 3)).cycle, 20);
}

// run() re-initialises: two runs give the same events and grid
LGD_TEST(serial_run_twice)
{
    std::vector<Event> ev1, ev2;
    SerialSimulator<ConwayCell> sim(new ExampleLifeInitializer(60));
    sim.addWriter(new RecordingWriter<ConwayCell>(&ev1, 7));
    sim.run();
    Grid<ConwayCell> first(sim.getGrid()->boundingBox());
    CoordBox<2> box = first.boundingBox();
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        first.set(*i, sim.getGrid()->get(*i));
    }
    sim.addWriter(new RecordingWriter<ConwayCell>(&ev2, 7));
    ev1.clear();
    sim.run();
    CHECK(ev1 == ev2);
    CHECK(*sim.getGrid() == first);
}

// the reference example's glider crosses the board: after 4k generations it
// sits k cells further along +x +y
LGD_TEST(serial_life_glider)
{
    SerialSimulator<ConwayCell> sim(new ExampleLifeInitializer(40));
    for (int i = 0; i < 40; ++i) {
        sim.step();
    }
    const int glider[5][2] = {{11, 10}, {12, 11}, {10, 12}, {11, 12}, {12, 12}};
    for (const auto& c : glider) {
        CHECK(sim.getGrid()->get(Coord<2>(c[0] + 10, c[1] + 10)).alive);
    }
}

// ---------------------------------------------------------------------
// SerialBOVWriter (io/test/unit/serialbovwritertest.h testBasic: period 4
// over 21 steps gives snapshots 0, 4, ..., 20, 21 with this header)
// ---------------------------------------------------------------------
static std::string slurp(const std::string& name)
{
    std::ifstream in(name.c_str(), std::ios::binary);
    std::stringstream buf;
    buf << in.rdbuf();
    return buf.str();
}

LGD_TEST(bov_writer_files)
{
    const Coord<3> dim(13, 12, 11);
    SerialSimulator<JacobiCell> sim(new JacobiInitializer(dim, 21, 5, 0.0));
    sim.addWriter(new SerialBOVWriter<JacobiCell>(Selector<JacobiCell>(&JacobiCell::temp, "val"), "lgdbovtest", 4));
    Grid<JacobiCell, Topologies::Cube<3>::Topology> initial(dim);
    JacobiInitializer(dim, 21, 5, 0.0).grid(&initial);
    sim.run();
    const unsigned steps[] = {0, 4, 8, 12, 16, 20, 21};
    for (unsigned step : steps) {
        char stem[64];
        std::snprintf(stem, sizeof(stem), "lgdbovtest.%05u", step);
        std::ostringstream exp;
        exp << "TIME: " << step << "\n"
            << "DATA_FILE: " << stem << ".data\n"
            << "DATA_SIZE: 13 12 11\n"
            << "DATA_FORMAT: DOUBLE\n"
            << "VARIABLE: val\n"
            << "DATA_ENDIAN: LITTLE\n"
            << "BRICK_ORIGIN: 0 0 0\n"
            << "BRICK_SIZE: 13 12 11\n"
            << "DIVIDE_BRICK: true\n"
            << "DATA_BRICKLETS: 13 12 11\n"
            << "DATA_COMPONENTS: 1\n";
        CHECK_EQ(slurp(std::string(stem) + ".bov"), exp.str());
        const std::string data = slurp(std::string(stem) + ".data");
        CHECK_EQ(data.size(), (std::size_t)dim.prod() * sizeof(double));
        if (step == 0 || step == 21) {
            const GridBase<JacobiCell, 3>& g =
                step == 0 ? static_cast<const GridBase<JacobiCell, 3>&>(initial) : *sim.getGrid();
            bool same = true;
            CoordBox<3> box(Coord<3>(), dim);
            for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
                double v;
                std::memcpy(&v, data.data() + i->toIndex(dim) * sizeof(double), sizeof(double));
                same = same && v == g.get(*i).temp;
            }
            CHECK(same);
        }
        std::remove((std::string(stem) + ".bov").c_str());
        std::remove((std::string(stem) + ".data").c_str());
    }
    // no snapshot off the period
    CHECK(slurp("lgdbovtest.00005.bov").empty());
}

LGD_TEST(bov_writer_2d_byte_and_type_names)
{
    SerialSimulator<ConwayCell> sim(new ExampleLifeInitializer(3));
    sim.addWriter(new SerialBOVWriter<ConwayCell>(&ConwayCell::alive, "lgdbovlife", 3));
    sim.run();
    const std::string h = slurp("lgdbovlife.00003.bov");
    CHECK(h.find("DATA_SIZE: 160 90 1\n") != std::string::npos);
    CHECK(h.find("DATA_FORMAT: BYTE\n") != std::string::npos);
    CHECK(h.find("VARIABLE: lgdbovlife\n") != std::string::npos);
    CHECK_EQ(slurp("lgdbovlife.00003.data").size(), (std::size_t)(160 * 90));
    for (unsigned step : {0u, 3u}) {
        char stem[64];
        std::snprintf(stem, sizeof(stem), "lgdbovlife.%05u", step);
        std::remove((std::string(stem) + ".bov").c_str());
        std::remove((std::string(stem) + ".data").c_str());
    }
    struct Odd { unsigned short u; };
    CHECK_THROWS(Selector<Odd>(&Odd::u, "u").typeName(), std::invalid_argument);
    CHECK_EQ(Selector<Odd>(&Odd::u, "u").sizeOfExternal(), sizeof(unsigned short));
}

// Torus topology in the oracle: a glider moves (1, 1) every 4 generations,
// so on a 16 x 16 torus it is back in place after 64 (and has crossed both
// seams); on the Cube it would have died against the edge
LGD_TEST(oracle_torus_glider_returns)
{
    class GliderInit : public SimpleInitializer<TorusLifeCell>
    {
    public:
        explicit GliderInit(unsigned steps) : SimpleInitializer<TorusLifeCell>(Coord<2>(16, 16), steps) {}
        virtual void grid(GridBase<TorusLifeCell, 2> *ret)
        {
            const int g[5][2] = {{13, 12}, {14, 13}, {12, 14}, {13, 14}, {14, 14}};
            for (auto& c : g) {
                ret->set(Coord<2>(c[0], c[1]), TorusLifeCell(true));
            }
        }
    };
    SerialSimulator<TorusLifeCell> at64(new GliderInit(64)), at0(new GliderInit(0)), at32(new GliderInit(32));
    at64.run();
    at0.run();
    at32.run();
    CHECK(*at64.getGrid() == *at0.getGrid());
    // halfway: moved by (8, 8), still five cells
    int alive = 0;
    bool moved = true;
    const int g[5][2] = {{13, 12}, {14, 13}, {12, 14}, {13, 14}, {14, 14}};
    for (auto& c : g) {
        moved = moved && at32.getGrid()->get(Coord<2>((c[0] + 8) % 16, (c[1] + 8) % 16)).alive;
    }
    CoordBox<2> box(Coord<2>(), Coord<2>(16, 16));
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        alive += at32.getGrid()->get(*i).alive;
    }
    CHECK(moved);
    CHECK_EQ(alive, 5);
}

// The probe cell of the generic-kernel GPU tests catches the mistakes it is
// there for (planted here, on the CPU): a wrong topology, a neighbour from
// another cycle, at the wrong place, or invalid, and a wrong nano step.
typedef ProbeCell<2, Topologies::Cube<2>::Topology, 1, Stencils::Moore<2, 1> > CubeProbe;
typedef ProbeCell<2, Topologies::Cube<2>::Topology, 1, Stencils::Moore<2, 1>, Topologies::Torus<2>::Topology>
    CubeProbeCheckedAsTorus;

template<typename CELL>
struct ProbeHood
{
    std::vector<CELL> cells;   // 3 x 3, x fastest
    const CELL& operator[](const Coord<2>& rel) const { return cells[(rel.y() + 1) * 3 + rel.x() + 1]; }
    template<int X, int Y, int Z>
    const CELL& operator[](FixedCoord<X, Y, Z>) const { return cells[(Y + 1) * 3 + X + 1]; }
};

static ProbeHood<CubeProbe> probeHood(const Coord<2>& center, const Coord<2>& dim)
{
    ProbeHood<CubeProbe> h;
    for (int y = -1; y < 2; ++y) {
        for (int x = -1; x < 2; ++x) {
            h.cells.push_back(CubeProbe(center + Coord<2>(x, y), dim, 0.5f));
        }
    }
    return h;
}

LGD_TEST(probe_cell_detects_planted_mistakes)
{
    const Coord<2> dim(10, 8);
    // consistent neighbourhood: valid, cycle advanced, value averaged
    {
        ProbeHood<CubeProbe> h = probeHood(Coord<2>(4, 4), dim);
        CubeProbe c;
        c.update(h, 0);
        CHECK(c.valid && c.bad == 0 && c.cycle == 1u && c.value == 0.5f);
    }
    // neighbour one cycle ahead / misplaced / invalid; wrong nano step
    const struct { int which; int bit; } cases[] = {
        {0, PROBE_CYCLE}, {1, PROBE_POSITION}, {2, PROBE_INVALID_NEIGHBOR}, {3, PROBE_NANO_STEP}};
    for (auto& k : cases) {
        ProbeHood<CubeProbe> h = probeHood(Coord<2>(4, 4), dim);
        unsigned nano = 0;
        if (k.which == 0) h.cells[2].cycle = 1;
        if (k.which == 1) h.cells[5].pos = Coord<2>(6, 4);
        if (k.which == 2) h.cells[7].valid = 0;
        if (k.which == 3) nano = 1;
        CubeProbe c;
        c.update(h, nano);
        CHECK(!c.valid);
        CHECK_EQ(c.bad, k.bit);
    }
    // the edge cell where the topology has none
    {
        ProbeHood<CubeProbe> h = probeHood(Coord<2>(0, 4), dim);
        h.cells[3] = CubeProbe::makeEdge(dim);
        h.cells[0] = CubeProbe::makeEdge(dim);
        h.cells[6] = CubeProbe::makeEdge(dim);
        CubeProbe c;
        c.update(h, 0);
        CHECK(c.valid);                                  // correct for the Cube
        h.cells[3] = CubeProbe(Coord<2>(-1, 4), dim, 0.5f);
        c = CubeProbe();
        c.update(h, 0);
        CHECK(!c.valid && (c.bad & PROBE_BAD_EDGE));     // a cell where the edge belongs
    }
    // a whole run whose checks assume the wrong topology: exactly the
    // boundary cells fail in the first step
    SerialSimulator<CubeProbeCheckedAsTorus> sim(new ProbeInitializer<CubeProbeCheckedAsTorus>(dim, 1, 3));
    sim.run();
    int invalid = 0, edgeBits = 0;
    CoordBox<2> box(Coord<2>(), dim);
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        const CubeProbeCheckedAsTorus c = sim.getGrid()->get(*i);
        invalid += !c.valid;
        edgeBits += (c.bad & PROBE_BAD_EDGE) != 0;
    }
    CHECK_EQ(invalid, 2 * (10 + 8) - 4);
    CHECK_EQ(edgeBits, invalid);
}

// ---------------------------------------------------------------------
// MPI-IO snapshots (io/mpiio.h:26-233 format, written by one process)
// ---------------------------------------------------------------------
static std::vector<char> fileBytes(const std::string& name)
{
    std::ifstream f(name.c_str(), std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

LGD_TEST(mpiio_writer_file_layout)
{
    const Coord<3> dim(7, 5, 3);
    SerialSimulator<JacobiCell> sim(new RandomCellInitializer<JacobiCell, 3>(
        dim, 7, 11, [](double u) { return JacobiCell(u); }, JacobiCell(-2.5)));
    sim.addWriter(new MPIIOWriter<JacobiCell>("snap", 3, 7));
    sim.run();
    // initial step, every 3rd step, last step
    for (unsigned step : {0u, 3u, 6u, 7u}) {
        char name[32];
        std::snprintf(name, sizeof(name), "snap%05u.mpiio", step);
        const std::vector<char> b = fileBytes(name);
        CHECK_EQ(b.size(), (std::size_t)(3 * 4 + 2 * 4 + 8 + 7 * 5 * 3 * 8));
        if (b.size() != (std::size_t)(3 * 4 + 2 * 4 + 8 + 7 * 5 * 3 * 8)) {
            continue;
        }
        int32_t dims[3];
        uint32_t st, mx;
        double edge;
        std::memcpy(dims, &b[0], 12);
        std::memcpy(&st, &b[12], 4);
        std::memcpy(&mx, &b[16], 4);
        std::memcpy(&edge, &b[20], 8);
        CHECK(dims[0] == 7 && dims[1] == 5 && dims[2] == 3);
        CHECK_EQ(st, step);
        CHECK_EQ(mx, 7u);
        CHECK_EQ(edge, -2.5);
    }
    // the last file's body is the final grid, x fastest
    const std::vector<char> b = fileBytes("snap00007.mpiio");
    const GridBase<JacobiCell, 3> *g = sim.getGrid();
    CoordBox<3> box(Coord<3>(), dim);
    int differ = 0;
    for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
        double v;
        std::memcpy(&v, &b[28 + 8 * i->toIndex(dim)], 8);
        differ += v != g->get(*i).temp;
    }
    CHECK_EQ(differ, 0);
}

LGD_TEST(mpiio_restart_matches_uninterrupted_run)
{
    const Coord<2> dim(61, 37);
    auto make = [&]() { return new RandomLifeInitializer(dim, 40, 5, 0.35); };
    std::vector<Event> evFull, evResumed;
    SerialSimulator<ConwayCell> full(make());
    full.addWriter(new MPIIOWriter<ConwayCell>("life", 10, 40));
    full.addWriter(new RecordingWriter<ConwayCell>(&evFull, 4));
    full.run();
    MPIIOInitializer<ConwayCell> *init = new MPIIOInitializer<ConwayCell>("life00020.mpiio");
    CHECK_EQ(init->startStep(), 20u);
    CHECK_EQ(init->maxSteps(), 40u);
    CHECK(init->gridDimensions() == dim);
    SerialSimulator<ConwayCell> resumed(init);
    resumed.addWriter(new RecordingWriter<ConwayCell>(&evResumed, 4));
    resumed.run();
    CHECK_EQ(resumed.getStep(), 40u);
    int differ = 0;
    CoordBox<2> box(Coord<2>(), dim);
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        differ += full.getGrid()->get(*i).alive != resumed.getGrid()->get(*i).alive;
    }
    CHECK_EQ(differ, 0);
    // the resumed writer keeps the phase: initialised at 20, then 24, 28, ...
    CHECK(!evResumed.empty() && evResumed.front().step == 20u && evResumed.front().event == WRITER_INITIALIZED);
    std::vector<Event> tail(evFull.end() - (long)(evResumed.size() - 1), evFull.end());
    std::vector<Event> resumedTail(evResumed.begin() + 1, evResumed.end());
    CHECK(tail == resumedTail);
}

LGD_TEST(mpiio_regions_and_torus_normalisation)
{
    typedef Topologies::Torus<2>::Topology Torus;
    const Coord<2> dim(9, 4);
    Grid<JacobiCell, Torus> grid(dim);
    grid.setEdge(JacobiCell(7.0));
    CoordBox<2> box(Coord<2>(), dim);
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        grid.set(*i, JacobiCell(i->x() + 100.0 * i->y()));
    }
    MPIIO<JacobiCell, Torus> io;
    Region<2> all;
    all << box;
    io.writeRegion(grid, dim, 3, 9, "torus.mpiio", all);
    // rewrite one row through a streak left of the box: on the torus it
    // lands on row 2, columns 7 and 8 (normalised origin), other bytes stay
    Grid<JacobiCell, Torus> changed = grid;
    changed.set(Coord<2>(7, 2), JacobiCell(-1.0));
    changed.set(Coord<2>(8, 2), JacobiCell(-2.0));
    Region<2> part;
    part << Streak<2>(Coord<2>(-2, 2), 0);
    io.writeRegion(changed, dim, 3, 9, "torus.mpiio", part);
    Grid<JacobiCell, Torus> back(dim);
    io.readRegion(&back, "torus.mpiio", all);
    CHECK_EQ(back.getEdge().temp, 7.0);
    int differ = 0;
    for (CoordBox<2>::Iterator i = box.begin(); i != box.end(); ++i) {
        differ += back.get(*i).temp != changed.get(*i).temp;
    }
    CHECK_EQ(differ, 0);
    CHECK_EQ(back.get(Coord<2>(7, 2)).temp, -1.0);
    Coord<2> d;
    unsigned step, maxSteps;
    io.readMetadata(&d, &step, &maxSteps, "torus.mpiio");
    CHECK(d == dim && step == 3u && maxSteps == 9u);
}

LGD_TEST(mpiio_errors)
{
    MPIIO<JacobiCell, Topologies::Cube<3>::Topology> io;
    Coord<3> d;
    unsigned s, m;
    CHECK_THROWS(io.readMetadata(&d, &s, &m, "no_such_file.mpiio"), std::runtime_error);
    {
        std::ofstream f("short.mpiio", std::ios::binary);
        const int32_t dims[3] = {4, 4, 4};
        f.write(reinterpret_cast<const char *>(dims), sizeof(dims));
    }
    CHECK_THROWS(io.readMetadata(&d, &s, &m, "short.mpiio"), std::runtime_error);
    CHECK_THROWS(MPIIOInitializer<JacobiCell>("short.mpiio"), std::runtime_error);
}

// ---------------------------------------------------------------------
// autotuner building blocks (misc/simulationparameters.h, optimizer.h,
// patternoptimizer.cpp, simulationfactory.h, io/varstepinitializerproxy.h)
// ---------------------------------------------------------------------
LGD_TEST(simulation_parameters_semantics)
{
    SimulationParameters p;
    p.addParameter("BlockDimX", 1, 128);
    p.addParameter("Coarse", 0, 64, 4.0);
    p.addParameter("Width", std::vector<int>{64, 128, 256});
    CHECK_EQ(p.size(), (std::size_t)3);
    CHECK_EQ((int)p["BlockDimX"], 1);
    p["BlockDimX"].setValue(10.7);                 // index rounds down to 10
    CHECK_EQ((int)p["BlockDimX"], 11);
    CHECK_EQ(p["BlockDimX"].getValue(), 10.0);
    p["BlockDimX"].setValue(500);                  // the interval is [1, 128)
    CHECK_EQ((int)p["BlockDimX"], 127);
    p["BlockDimX"].setValue(-3);
    CHECK_EQ((int)p["BlockDimX"], 1);
    p["BlockDimX"] += 5;
    CHECK_EQ((int)p["BlockDimX"], 6);
    p["Coarse"].setValue(10);                      // multiples of the granularity
    CHECK_EQ((int)p["Coarse"], 8);
    p["Coarse"] += 3;
    CHECK_EQ((int)p["Coarse"], 8);
    p["Coarse"] += 100;
    CHECK_EQ((int)p["Coarse"], 60);
    CHECK_EQ((int)p["Width"], 64);
    p["Width"].setValue(1.9);
    CHECK(p["Width"] == 128);
    p["Width"] += 10;
    CHECK_EQ((int)p["Width"], 256);
    CHECK_EQ(p["Width"].getMax(), 3.0);
    CHECK_THROWS((void)(double)p["Width"], std::logic_error);
    CHECK_THROWS(p["NoSuch"], std::invalid_argument);
    SimulationParameters copy(p);                  // deep copy
    copy["Width"].setValue(0);
    CHECK_EQ((int)p["Width"], 256);
    CHECK_EQ((int)copy["Width"], 64);
    p.replaceParameter("Width", std::vector<int>{1, 2});
    CHECK_EQ((int)p["Width"], 1);
}

class Paraboloid : public Optimizer::Evaluator
{
public:
    int calls = 0;
    double operator()(const SimulationParameters& p)
    {
        ++calls;
        const double x = (int)p["x"], y = (int)p["y"], z = (int)p["z"];
        return -((x - 37) * (x - 37) + 3 * (y - 5) * (y - 5) + 10 * (z - 128) * (z - 128) / 4096.0);
    }
};

LGD_TEST(pattern_optimizer_finds_maximum)
{
    SimulationParameters p;
    p.addParameter("x", 0, 100);
    p.addParameter("y", 0, 16);
    p.addParameter("z", std::vector<int>{16, 32, 64, 128, 256, 512});
    PatternOptimizer opt(p);
    Paraboloid f;
    SimulationParameters best = opt(200, f);
    CHECK_EQ((int)best["x"], 37);
    CHECK_EQ((int)best["y"], 5);
    CHECK_EQ((int)best["z"], 128);
    CHECK_EQ(opt.getFitness(), 0.0);
    CHECK(f.calls < 100 * 16 * 6 / 20);           // far fewer than the 9600 points
    // explicit step widths
    PatternOptimizer small(p, std::vector<double>{1, 1, 1});
    Paraboloid g;
    best = small(500, g);
    CHECK(best["x"] == 37 && best["y"] == 5 && best["z"] == 128);
    CHECK_THROWS(PatternOptimizer(p, std::vector<double>{1, 1}), std::invalid_argument);
}

LGD_TEST(var_step_initializer_proxy)
{
    VarStepInitializerProxy<ConwayCell> proxy(new RandomLifeInitializer(Coord<2>(20, 10), 50, 1, 0.3, 7));
    CHECK_EQ(proxy.startStep(), 7u);
    CHECK_EQ(proxy.maxSteps(), 50u);
    proxy.setMaxSteps(3);
    CHECK_EQ(proxy.maxSteps(), 10u);
    std::unique_ptr<ClonableInitializer<ConwayCell> > clone(proxy.clone());
    CHECK_EQ(clone->maxSteps(), 10u);
    proxy.resetMaxSteps();
    CHECK_EQ(proxy.maxSteps(), 50u);
    Grid<ConwayCell> a(Coord<2>(20, 10)), b(Coord<2>(20, 10));
    proxy.grid(&a);
    clone->grid(&b);
    CHECK(a == b);
}

// a test-only factory over the oracle simulator: the factory mechanics
// (fitness is minus the run time, failures rate worst, plugins cloned in)
class OracleLifeFactory : public SimulationFactory<ConwayCell>
{
public:
    explicit OracleLifeFactory(InitPtr init) : SimulationFactory<ConwayCell>(init)
    {
        parameterSet.addParameter("Fail", std::vector<int>{0, 1});
    }
    std::string name() const { return "oracle"; }

protected:
    Simulator<ConwayCell> *buildSimulator(InitPtr init, const SimulationParameters& p) const
    {
        if ((int)p["Fail"]) {
            throw std::runtime_error("configuration rejected");
        }
        SerialSimulator<ConwayCell> *sim = new SerialSimulator<ConwayCell>(init->clone());
        addPlugins(sim);
        return sim;
    }
};

LGD_TEST(simulation_factory_mechanics)
{
    std::shared_ptr<VarStepInitializerProxy<ConwayCell> > proxy(
        new VarStepInitializerProxy<ConwayCell>(new RandomLifeInitializer(Coord<2>(40, 30), 12, 3)));
    OracleLifeFactory factory(proxy);
    std::vector<Event> events;
    factory.addWriter(RecordingWriter<ConwayCell>(&events, 4));
    SimulationParameters p = factory.parameters();
    const double ok = factory(p);
    CHECK(ok < 0 && ok > -10);
    CHECK_EQ(events.size(), (std::size_t)4);        // 0, 4, 8, 12 through the cloned writer
    p["Fail"].setValue(1);
    CHECK_EQ(factory(p), std::numeric_limits<double>::lowest());
    std::unique_ptr<Simulator<ConwayCell> > sim(factory());
    sim->run();
    CHECK_EQ(sim->getStep(), 12u);
}

// ---------------------------------------------------------------------
// struct-of-arrays storage and the oracle's SoA / line update paths
// ---------------------------------------------------------------------

// the registration's addressing: member m of element i at
// data + DIM_PROD * offset(m) + i * sizeof(m), offsets the summed sizes of
// the members before (libflatarray detail/macros.hpp:99-106)
LGD_TEST(soa_registration_member_addresses)
{
    typedef LibFlatArray::soa_traits<SoAMixedCell> T;
    CHECK_EQ(T::MEMBERS, 3);
    CHECK_EQ(T::BYTES_PER_CELL, (std::size_t)(3 * 4 + 1 + 8));
    std::vector<std::size_t> offsets, sizes, counts;
    T::each_member([&](std::size_t o, std::size_t s, std::size_t c) {
        offsets.push_back(o);
        sizes.push_back(s);
        counts.push_back(c);
    });
    CHECK(offsets == std::vector<std::size_t>({0, 12, 13}));
    CHECK(sizes == std::vector<std::size_t>({4, 1, 8}));
    CHECK(counts == std::vector<std::size_t>({3, 1, 1}));
    const long prod = 40;
    std::vector<char> buf(prod * T::BYTES_PER_CELL, 0);
    LibFlatArray::soa_accessor<SoAMixedCell> acc(buf.data(), 7, 5, 20, prod);
    acc << SoAMixedCell(1.5f, true, 2.25);
    float f1;
    std::memcpy(&f1, buf.data() + prod * (0 + 1 * 4) + 7 * 4, 4);   // f[1], the second plane of f
    CHECK_EQ(f1, 3.0f);
    CHECK_EQ(buf[prod * 12 + 7], (char)1);                          // on
    double d;
    std::memcpy(&d, buf.data() + prod * 13 + 7 * 8, 8);
    CHECK_EQ(d, 2.25);
    const SoAMixedCell back = acc;
    CHECK(back == SoAMixedCell(1.5f, true, 2.25));
    // coord<X, Y, Z> moves the index by X + Y DIM_X + Z DIM_X DIM_Y
    const long moved = acc[LibFlatArray::coord<1, -1, 1>()].index();
    CHECK_EQ(moved, 7L + 1 - 5 + 20);
}

// edge radii only on unwrapped axes (soagrid.h:578-584); the padding holds
// the edge cell; get / set through the topology
LGD_TEST(soa_grid_edges_and_topology)
{
    typedef SoAGrid<SoAJacobiCell, Topologies::Cube<3>::Topology> CubeGrid;
    typedef SoAGrid<SoATorusJacobiCell, Topologies::Torus<3>::Topology> TorusGrid;
    typedef TopologiesHelpers::Topology<3, true, false, true> Slab;
    typedef SoAGrid<SoASlabJacobiCell, Slab> SlabGrid;
    CHECK(CubeGrid::calcEdgeRadii() == Coord<3>(1, 1, 1));
    CHECK(TorusGrid::calcEdgeRadii() == Coord<3>(0, 0, 0));
    CHECK(SlabGrid::calcEdgeRadii() == Coord<3>(0, 1, 0));
    SoAGrid<SoAJacobiCell, Topologies::Cube<3>::Topology> g(CoordBox<3>(Coord<3>(), Coord<3>(4, 3, 2)),
                                                            SoAJacobiCell(1.0), SoAJacobiCell(-5.0));
    CHECK(g.soaLayout().dimX == 6 && g.soaLayout().dimY == 5 && g.soaLayout().dimZ == 4);
    g.set(Coord<3>(3, 2, 1), SoAJacobiCell(7.0));
    CHECK_EQ(g.get(Coord<3>(3, 2, 1)).temp, 7.0);
    CHECK_EQ(g.get(Coord<3>(4, 2, 1)).temp, -5.0);
    // the padding element right of (3, 2, 1) holds the edge cell
    CHECK_EQ(g.accessor(g.soaLayout().index(4, 2, 1)).temp(), -5.0);
    g.setEdge(SoAJacobiCell(9.0));
    CHECK_EQ(g.accessor(g.soaLayout().index(-1, 0, 0)).temp(), 9.0);
    CHECK_EQ(g.get(Coord<3>(0, 0, 0)).temp, 1.0);
    SoAGrid<SoATorusJacobiCell, Topologies::Torus<3>::Topology> t(CoordBox<3>(Coord<3>(), Coord<3>(4, 3, 2)));
    t.set(Coord<3>(-1, 3, 2), SoATorusJacobiCell(4.0));     // wraps to (3, 0, 0)
    CHECK_EQ(t.get(Coord<3>(3, 0, 0)).temp, 4.0);
}

// the reference functor test's closed forms on the oracle's SoA path
template<typename CELL>
int oracleFunctorMismatches(const Coord<3>& dim, double edgeA, bool torus)
{
    SerialSimulator<CELL> sim(new EncodedInitializer<CELL>(dim, 1, edgeA));
    sim.run();
    const GridBase<CELL, 3> *g = sim.getGrid();
    int bad = 0;
    CoordBox<3> box(Coord<3>(), dim);
    const int dx[6] = {0, 0, -1, 1, 0, 0}, dy[6] = {0, -1, 0, 0, 1, 0}, dz[6] = {-1, 0, 0, 0, 0, 1};
    for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
        Coord<3> sum;
        for (int k = 0; k < 6; ++k) {
            Coord<3> n = *i + Coord<3>(dx[k], dy[k], dz[k]);
            if (torus) {
                n = Topologies::Torus<3>::Topology::normalize(n, dim);
            } else if (Topologies::Cube<3>::Topology::isOutOfBounds(n, dim)) {
                n = Coord<3>(-1, -1, -1);
            }
            sum = sum + n;
        }
        bad += g->get(*i).valB != sum.x() + sum.y() * 1000.0 + sum.z() * 1000.0 * 1000.0;
    }
    return bad;
}

LGD_TEST(oracle_soa_functor_known_answers)
{
    CHECK_EQ(oracleFunctorMismatches<TwoDoublesTorus>(Coord<3>(200, 100, 50), -1.0, true), 0);
    CHECK_EQ(oracleFunctorMismatches<TwoDoublesTorus>(Coord<3>(1, 10, 20), -1.0, true), 0);
    CHECK_EQ(oracleFunctorMismatches<TwoDoublesCube>(Coord<3>(200, 100, 50), -1001001.0, false), 0);
}

// the three update paths of the oracle agree on the same arithmetic: SoA
// streaks, array-of-structs lines and cell-wise update()
template<typename A, typename B>
int jacobiDifferences(const Coord<3>& dim, unsigned steps, double edge)
{
    SerialSimulator<A> a(new RandomCellInitializer<A, 3>(dim, steps, 9, [](double u) { return A(u); }, A(edge)));
    SerialSimulator<B> b(new RandomCellInitializer<B, 3>(dim, steps, 9, [](double u) { return B(u); }, B(edge)));
    a.run();
    b.run();
    int differ = 0;
    CoordBox<3> box(Coord<3>(), dim);
    for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
        differ += a.getGrid()->get(*i).temp != b.getGrid()->get(*i).temp;
    }
    return differ;
}

LGD_TEST(oracle_soa_and_line_paths_equal_cellwise_update)
{
    static_assert(SerialSimulator<SoAJacobiCell>::SOA && !SerialSimulator<SoAJacobiCell>::LINE, "SoA path");
    static_assert(SerialSimulator<LineJacobiCell>::LINE && !SerialSimulator<LineJacobiCell>::SOA, "line path");
    static_assert(!SerialSimulator<JacobiCell>::SOA && !SerialSimulator<JacobiCell>::LINE, "cell-wise path");
    CHECK_EQ((jacobiDifferences<SoAJacobiCell, JacobiCell>(Coord<3>(23, 17, 11), 6, 0.25)), 0);
    CHECK_EQ((jacobiDifferences<LineJacobiCell, JacobiCell>(Coord<3>(23, 17, 11), 6, 0.25)), 0);
    CHECK_EQ((jacobiDifferences<SoATorusJacobiCell, TorusJacobiCell>(Coord<3>(9, 7, 5), 5, 0.0)), 0);
    CHECK_EQ((jacobiDifferences<SoASlabJacobiCell, SlabJacobiCell>(Coord<3>(9, 7, 5), 5, 0.5)), 0);
    // one-cell axes: the first and last cell are the same
    CHECK_EQ((jacobiDifferences<SoATorusJacobiCell, TorusJacobiCell>(Coord<3>(1, 3, 1), 4, 0.0)), 0);
    CHECK_EQ((jacobiDifferences<SoAJacobiCell, JacobiCell>(Coord<3>(1, 1, 1), 3, 0.5)), 0);
}

// SoA probe cells pass their own checks on the oracle, and a probe built
// to assume the wrong topology fails them (the checks see the wrap)
template<typename CELL>
int oracleProbeInvalid(const Coord<CELL::API::Topology::DIM>& dim, unsigned steps)
{
    SerialSimulator<CELL> sim(new ProbeInitializer<CELL>(dim, steps, 3));
    sim.run();
    int invalid = 0;
    CoordBox<CELL::API::Topology::DIM> box(Coord<CELL::API::Topology::DIM>(), dim);
    for (typename CoordBox<CELL::API::Topology::DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
        const CELL c = sim.getGrid()->get(*i);
        invalid += !c.valid || c.cycle != steps * APITraits::SelectNanoSteps<CELL>::VALUE;
    }
    return invalid;
}

LGD_TEST(oracle_soa_probe_cells)
{
    CHECK_EQ(oracleProbeInvalid<SoAProbe3Cube>(Coord<3>(9, 7, 5), 3), 0);
    CHECK_EQ(oracleProbeInvalid<SoAProbe3Torus>(Coord<3>(9, 7, 5), 3), 0);
    CHECK_EQ(oracleProbeInvalid<SoAProbe3Tube>(Coord<3>(9, 7, 5), 3), 0);
    CHECK_EQ(oracleProbeInvalid<SoAProbe2Torus>(Coord<2>(11, 6), 4), 0);
    CHECK_EQ(oracleProbeInvalid<SoAProbe1Ring>(Coord<1>(13), 4), 0);
    CHECK_EQ(oracleProbeInvalid<SoAProbe3Torus>(Coord<3>(1, 2, 1), 2), 0);
}

int main(int argc, char **argv)
{
    return lgdtest::runAll(argc, argv);
}

// ---------------------------------------------------------------------
// TrackingGrid's on-demand reads (the host side of HipSimulator's partial
// device-to-host transfers for writers), with a host array as the device
// ---------------------------------------------------------------------
LGD_TEST(tracking_grid_reads_rows_on_demand)
{
    typedef Topologies::Cube<3>::Topology Cube3;
    const CoordBox<3> box(Coord<3>(2, 3, 4), Coord<3>(10, 40, 30));   // 1200 rows, one-by-one limit 37
    auto truth = [](const Coord<3>& c) { return 1000.0 * c.z() + 10.0 * c.y() + 0.5 * c.x(); };
    TrackingGrid<double, Cube3> g(box);
    g.setEdge(-1.0);
    std::vector<Region<3> > calls;
    int alls = 0;
    auto fill = [&](const Streak<3>& s) {
        for (Coord<3> c = s.origin; c.x() < s.endX; ++c.x()) {
            g.data()[(c - box.origin).toIndex(box.dimensions)] = truth(c);
        }
    };
    g.markStale([&](const Region<3>& r) {
                    calls.push_back(r);
                    for (const Streak<3>& s : r.streaks()) {
                        fill(s);
                    }
                },
                [&]() {
                    ++alls;
                    for (int z = 0; z < 30; ++z) {
                        for (int y = 0; y < 40; ++y) {
                            fill(Streak<3>(Coord<3>(2, 3 + y, 4 + z), 12));
                        }
                    }
                });
    CHECK(g.hasStaleRows());
    CHECK_EQ(g.rowLimit(), (std::size_t)37);
    // the first read of a row brings in exactly that row
    CHECK_EQ(g.get(Coord<3>(5, 13, 11)), truth(Coord<3>(5, 13, 11)));
    CHECK_EQ(calls.size(), (std::size_t)1);
    Region<3> row;
    row << Streak<3>(Coord<3>(2, 13, 11), 12);
    CHECK(calls[0] == row);
    CHECK_EQ(g.get(Coord<3>(11, 13, 11)), truth(Coord<3>(11, 13, 11)));
    CHECK_EQ(calls.size(), (std::size_t)1);
    // edge reads never fetch
    CHECK_EQ(g.get(Coord<3>(1, 13, 11)), -1.0);
    CHECK_EQ(g.get(Coord<3>(5, 43, 11)), -1.0);
    CHECK_EQ(calls.size(), (std::size_t)1);
    // a streak read fetches its row once
    double line[4];
    g.get(Streak<3>(Coord<3>(4, 20, 33), 8), line);
    CHECK_EQ(line[3], truth(Coord<3>(7, 20, 33)));
    CHECK_EQ(calls.size(), (std::size_t)2);
    // past the limit the rest comes in one transfer
    for (int y = 3; calls.size() < 37; ++y) {
        g.get(Coord<3>(2, y, 4));
    }
    CHECK_EQ(alls, 0);
    CHECK(g.hasStaleRows());
    CHECK_EQ(g.get(Coord<3>(3, 42, 33)), truth(Coord<3>(3, 42, 33)));
    CHECK_EQ(alls, 1);
    CHECK(!g.hasStaleRows());
    CHECK_EQ(g.rowsFetchedOnDemand(), (std::size_t)37);
    // after markCurrent nothing is fetched, and markStale starts a new count
    g.markCurrent();
    g.get(Coord<3>(2, 3, 4));
    CHECK_EQ(calls.size(), (std::size_t)37);
    bool fetched = false;
    g.markStale([&](const Region<3>&) { fetched = true; }, [&]() {});
    g.get(Coord<3>(2, 3, 4));
    CHECK(fetched);
}

LGD_TEST(tracking_grid_on_demand_torus_rows)
{
    typedef Topologies::Torus<2>::Topology Torus2;
    TrackingGrid<int, Torus2> g(CoordBox<2>(Coord<2>(), Coord<2>(7, 5)));
    std::vector<int> rows;
    g.markStale([&](const Region<2>& r) { rows.push_back(r.streaks()[0].origin.y()); }, [&]() {});
    g.get(Coord<2>(-1, -1));   // wraps to (6, 4)
    g.get(Coord<2>(3, 9));     // wraps to (3, 4): same row
    g.get(Coord<2>(0, 5));     // wraps to (0, 0)
    CHECK(rows == std::vector<int>({4, 0}));
}

LGD_TEST(tracing_writer_output_and_rates)
{
    typedef TracingWriter<ConwayCell> TW;
    CHECK_EQ(TW::renderDuration(3661.5), std::string("01:01:01.500000"));
    CHECK_EQ(TW::renderDuration(59.0), std::string("00:00:59"));
    // GLUPS = steps * cells / seconds / 1e9; bandwidth = GLUPS * 2 * sizeof(cell)
    CHECK(std::fabs(TW::glups(100, 1e6, 0.5) - 0.2) < 1e-12);
    CHECK(std::fabs(TW::bandwidthGBs(0.2) - 0.4 * sizeof(ConwayCell)) < 1e-12);
    CHECK_THROWS(TW(0, 10), std::invalid_argument);
    std::ostringstream out;
    SerialSimulator<ConwayCell> sim(new ExampleLifeInitializer(20));
    sim.addWriter(new TW(5, 20, TW::OUTPUT_ON_ALL_RANKS, out));
    sim.run();
    const std::string log = out.str();
    std::size_t steps = 0;
    for (std::size_t p = log.find("TracingWriter::stepFinished()"); p != std::string::npos;
         p = log.find("TracingWriter::stepFinished()", p + 1)) {
        ++steps;
    }
    CHECK_EQ(steps, (std::size_t)3);   // 5, 10, 15; step 20 is ALL_DONE
    CHECK(log.find("step: 15 of 20") != std::string::npos);
    CHECK(log.rfind("TracingWriter::allDone()") > log.rfind("stepFinished"));
    // a writer meant for another rank prints nothing
    std::ostringstream none;
    SerialSimulator<ConwayCell> sim2(new ExampleLifeInitializer(10));
    sim2.addWriter(new TW(5, 10, 3, none));
    sim2.run();
    CHECK(none.str().empty());
}
