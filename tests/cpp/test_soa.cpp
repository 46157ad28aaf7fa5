// GPU tests of the struct-of-arrays path: cells registered with
// LIBFLATARRAY_REGISTER_SOA and updated by their own static
// updateLineX(hoodOld, indexEnd, hoodNew, nanoStep), run by HipSimulator on
// device SoA grids (libgeodecomp/hip/soafamily.h) and by the oracle
// SerialSimulator on host SoAGrids through the reference's streak functor
// semantics.  Grids at every writer callback and event sequences must agree
// exactly; the scalar SoA Jacobi cell must also equal the hand-written
// Jacobi family, and the functor cells the reference's closed forms
// (storage/test/unit/fixedneighborhoodupdatefunctortest.h).  The probe
// cases mirror the reference's TestCellSoA runs
// (parallelization/test/unit/cudasimulatortest.h:46-60, misc/testcell.h).
// The array-of-structs line cell (updateLineX(target, &x, endX, ...)) runs
// through the generic family's line path.
//
// Compiled with hipcc (tests/cpp/Makefile).
#include "compare.h"
#include "harness.h"
#include "soacells.h"

#include <libgeodecomp/hip/genericfamily.h>
#include <libgeodecomp/parallelization/hipsimulator.h>
#include <serialsimulator.h>

#include <cstdio>

typedef Topologies::Cube<3>::Topology Cube3;
typedef Topologies::Torus<3>::Topology Torus3;

static_assert(std::is_same<HipSimulator<SoAJacobiCell>::Family, HipKernels::SoA<SoAJacobiCell> >::value,
              "SoA line cells run through the SoA family");
static_assert(std::is_same<HipSimulator<LineJacobiCell>::Family, HipKernels::Generic<LineJacobiCell> >::value,
              "array-of-structs line cells run through the generic family");

template<typename CELL>
std::vector<CELL> finalGrid(HipSimulator<CELL>& sim)
{
    std::vector<CELL> out;
    const GridBase<CELL, 3> *g = sim.getGrid();
    CoordBox<3> box = g->boundingBox();
    for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
        out.push_back(g->get(*i));
    }
    return out;
}

// ---------------------------------------------------------------------
// scalar SoA Jacobi: oracle, and the hand-written family, bit for bit
// ---------------------------------------------------------------------
template<typename SOACELL, typename FAMILYCELL>
void soaJacobiCase(const Coord<3>& dim, unsigned steps, unsigned period, double edge)
{
    const uint64_t seed = 17;
    Snapshots<SOACELL> sh, sr;
    runBoth<SOACELL>(
        [&]() {
            return new RandomCellInitializer<SOACELL, 3>(dim, steps, seed, [](double u) { return SOACELL(u); },
                                                         SOACELL(edge));
        },
        period, Coord<3>(), &sh, &sr);
    HipSimulator<FAMILYCELL> family(new RandomCellInitializer<FAMILYCELL, 3>(
        dim, steps, seed, [](double u) { return FAMILYCELL(u); }, FAMILYCELL(edge)));
    family.run();
    const std::vector<FAMILYCELL> f = finalGrid(family);
    CHECK_EQ(f.size(), sh.grids.back().size());
    int differ = 0;
    for (std::size_t k = 0; k < f.size() && k < sh.grids.back().size(); ++k) {
        differ += f[k].temp != sh.grids.back()[k].temp;
    }
    CHECK_EQ(differ, 0);
}

LGD_TEST(soa_jacobi_cube_equals_oracle_and_jacobi_family)
{
    soaJacobiCase<SoAJacobiCell, JacobiCell>(Coord<3>(37, 29, 23), 12, 4, 0.25);
    soaJacobiCase<SoAJacobiCell, JacobiCell>(Coord<3>(64, 64, 64), 20, 5, 0.0);
}

LGD_TEST(soa_jacobi_torus_equals_oracle_and_jacobi_family)
{
    soaJacobiCase<SoATorusJacobiCell, TorusJacobiCell>(Coord<3>(33, 21, 17), 11, 3, 0.5);
}

LGD_TEST(soa_jacobi_slab_equals_oracle_and_jacobi_family)
{
    soaJacobiCase<SoASlabJacobiCell, SlabJacobiCell>(Coord<3>(19, 23, 13), 9, 3, 0.75);
}

// the time loop replays captured graphs: 100 quiet steps, one writer at the
// end, against the oracle
LGD_TEST(soa_jacobi_graph_replays)
{
    const Coord<3> dim(40, 24, 16);
    Snapshots<SoAJacobiCell> sh, sr;
    runBoth<SoAJacobiCell>(
        [&]() {
            return new RandomCellInitializer<SoAJacobiCell, 3>(dim, 100, 3,
                                                               [](double u) { return SoAJacobiCell(u); },
                                                               SoAJacobiCell(0.0));
        },
        25, Coord<3>(), &sh, &sr);
    HipSimulator<SoAJacobiCell> sim(new RandomCellInitializer<SoAJacobiCell, 3>(
        dim, 100, 3, [](double u) { return SoAJacobiCell(u); }, SoAJacobiCell(0.0)));
    std::vector<Event> events;
    sim.addWriter(new RecordingWriter<SoAJacobiCell>(&events, 25));
    sim.run();
    std::printf("  graphs captured %llu, graph launches %llu\n", sim.deviceFamily().graphs().graphsCaptured(),
                sim.deviceFamily().graphs().graphLaunches());
    CHECK(sim.deviceFamily().graphs().graphsCaptured() >= 1);
    CHECK(sim.deviceFamily().graphs().graphLaunches() >= 40);
}

// ---------------------------------------------------------------------
// the reference functor test's closed forms (one step):
// valB = sum of the six face neighbours' encodings x + 1000 y + 10^6 z,
// wrapped on the torus, the edge encoding (-1, -1, -1) on the cube
// ---------------------------------------------------------------------
template<typename CELL>
int functorMismatches(HipSimulator<CELL>& sim, const Coord<3>& dim, bool torus)
{
    const GridBase<CELL, 3> *g = sim.getGrid();
    int bad = 0;
    CoordBox<3> box(Coord<3>(), dim);
    const int dx[6] = {0, 0, -1, 1, 0, 0}, dy[6] = {0, -1, 0, 0, 1, 0}, dz[6] = {-1, 0, 0, 0, 0, 1};
    for (CoordBox<3>::Iterator i = box.begin(); i != box.end(); ++i) {
        Coord<3> sum;
        for (int k = 0; k < 6; ++k) {
            Coord<3> n = *i + Coord<3>(dx[k], dy[k], dz[k]);
            if (torus) {
                n = Torus3::normalize(n, dim);
            } else if (Cube3::isOutOfBounds(n, dim)) {
                n = Coord<3>(-1, -1, -1);
            }
            sum = sum + n;
        }
        const double expected = sum.x() + sum.y() * 1000.0 + sum.z() * 1000.0 * 1000.0;
        const CELL c = g->get(*i);
        bad += c.valB != expected || c.valA != i->x() + i->y() * 1000.0 + i->z() * 1000.0 * 1000.0;
    }
    return bad;
}

LGD_TEST(soa_functor_torus_known_answer)
{
    const Coord<3> dim(200, 100, 50);
    HipSimulator<TwoDoublesTorus> sim(new EncodedInitializer<TwoDoublesTorus>(dim, 1, -1.0));
    sim.run();
    CHECK_EQ(functorMismatches(sim, dim, true), 0);
}

LGD_TEST(soa_functor_degenerated_torus_known_answer)
{
    const Coord<3> dim(1, 10, 20);
    HipSimulator<TwoDoublesTorus> sim(new EncodedInitializer<TwoDoublesTorus>(dim, 1, -1.0));
    sim.run();
    CHECK_EQ(functorMismatches(sim, dim, true), 0);
}

LGD_TEST(soa_functor_cube_known_answer)
{
    const Coord<3> dim(200, 100, 50);
    HipSimulator<TwoDoublesCube> sim(new EncodedInitializer<TwoDoublesCube>(dim, 1, -1001001.0));
    sim.run();
    CHECK_EQ(functorMismatches(sim, dim, false), 0);
}

// ---------------------------------------------------------------------
// SoA probe cells (TestCellSoA): Cube, Torus, a tube (y and z wrapped),
// 1D / 2D / 3D, nano steps
// ---------------------------------------------------------------------
LGD_TEST(soa_probe_3d_cube_50x20x10)
{
    probeCase<SoAProbe3Cube>(Coord<3>(50, 20, 10), 7, 2);
}

LGD_TEST(soa_probe_3d_torus_50x20x10_three_nano_steps)
{
    probeCase<SoAProbe3Torus>(Coord<3>(50, 20, 10), 7, 2);
}

LGD_TEST(soa_probe_3d_tube_two_nano_steps)
{
    probeCase<SoAProbe3Tube>(Coord<3>(31, 17, 9), 5, 3);
}

LGD_TEST(soa_probe_2d_cube_121x151)
{
    probeCase<SoAProbe2Cube>(Coord<2>(121, 151), 11, 5);
}

LGD_TEST(soa_probe_2d_torus_two_nano_steps)
{
    probeCase<SoAProbe2Torus>(Coord<2>(121, 151), 11, 5);
}

LGD_TEST(soa_probe_1d_ring_777)
{
    probeCase<SoAProbe1Ring>(Coord<1>(777), 13, 4);
}

// axes of one and two cells: both wrap offsets of a cell apply at once
LGD_TEST(soa_probe_tiny_tori)
{
    probeCase<SoAProbe3Torus>(Coord<3>(1, 2, 1), 4, 1);
    probeCase<SoAProbe2Torus>(Coord<2>(2, 1), 6, 2);
    probeCase<SoAProbe1Ring>(Coord<1>(1), 3, 1);
}

LGD_TEST(soa_probe_block_shapes)
{
    probeCase<SoAProbe3Torus>(Coord<3>(70, 9, 5), 4, 2, Coord<3>(128, 2, 1));
    probeCase<SoAProbe3Torus>(Coord<3>(70, 9, 5), 4, 2, Coord<3>(256, 1, 1));
    CHECK_THROWS(HipSimulator<SoAProbe2Cube>(new ProbeInitializer<SoAProbe2Cube>(Coord<2>(8, 8), 1, 1),
                                             Coord<3>(64, 8, 1)),
                 std::invalid_argument);
}

// steerer edits reach the SoA device grids (region uploads per member plane)
static void nudgeSoA(GridBase<SoAProbe2Torus, 2> *grid, unsigned step, Steerer<SoAProbe2Torus>::SteererFeedback *,
                     void *)
{
    const Coord<2> c((int)(step * 13) % 40, (int)(step * 7) % 30);
    SoAProbe2Torus cell = grid->get(c);
    cell.value += 1.0f;
    grid->set(c, cell);
}

LGD_TEST(soa_probe_steerer_edits_uploaded)
{
    Snapshots<SoAProbe2Torus> sh, sr;
    runBoth<SoAProbe2Torus>([]() { return new ProbeInitializer<SoAProbe2Torus>(Coord<2>(45, 33), 12, 5); }, 3,
                            Coord<3>(), &sh, &sr, 2, &nudgeSoA);
    checkProbes(sh, 12, 2);
}

// ---------------------------------------------------------------------
// private members, an array member, mixed member sizes
// ---------------------------------------------------------------------
LGD_TEST(soa_mixed_members_equal_oracle)
{
    const Coord<2> dim(33, 17);
    Snapshots<SoAMixedCell> sh, sr;
    runBoth<SoAMixedCell>(
        [&]() {
            return new RandomCellInitializer<SoAMixedCell, 2>(
                dim, 9, 7, [](double u) { return SoAMixedCell((float)u, u < 0.5, u * 3.0); }, SoAMixedCell(-1.0f));
        },
        2, Coord<3>(), &sh, &sr);
}

// ---------------------------------------------------------------------
// array-of-structs line cell: the generic family's line path
// ---------------------------------------------------------------------
LGD_TEST(aos_line_jacobi_equals_oracle_and_jacobi_family)
{
    const Coord<3> dim(37, 29, 23);
    Snapshots<LineJacobiCell> sh, sr;
    runBoth<LineJacobiCell>(
        [&]() {
            return new RandomCellInitializer<LineJacobiCell, 3>(dim, 12, 17,
                                                                [](double u) { return LineJacobiCell(u); },
                                                                LineJacobiCell(0.25));
        },
        4, Coord<3>(), &sh, &sr);
    HipSimulator<JacobiCell> family(new RandomCellInitializer<JacobiCell, 3>(
        dim, 12, 17, [](double u) { return JacobiCell(u); }, JacobiCell(0.25)));
    family.run();
    const std::vector<JacobiCell> f = finalGrid(family);
    int differ = 0;
    for (std::size_t k = 0; k < f.size(); ++k) {
        differ += f[k].temp != sh.grids.back()[k].temp;
    }
    CHECK_EQ(differ, 0);
}

int main(int argc, char **argv)
{
    return lgdtest::runAll(argc, argv);
}
