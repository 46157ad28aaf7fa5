// HipSimulator<CELL>: single-GPU (MI355X, gfx950) drop-in for the
// reference's CUDASimulator step loop
// (src/libgeodecomp/parallelization/cudasimulator.h:297-596).
//
// Same user-facing contract as CUDASimulator/SerialSimulator: constructed
// from an Initializer it owns, Writers and Steerers added with
// addWriter()/addSteerer(), run()/step()/getGrid()/getStep()/
// gatherStatistics().  Event order follows SerialSimulator
// (serialsimulator.h:76-117, :144-180): STEERER_INITIALIZED and
// WRITER_INITIALIZED at the start of run(), steerers before every step whose
// number is a multiple of their period, writers after steps that are a
// multiple of their period and at the last step, STEERER_ALL_DONE at the
// end.
//
// The grid lives on the device between plugin callbacks.  The sweeps are
// the hand-written kernels of the cell's registered kernel family
// (libgeodecomp/hip/kernelfamilies.h), reached through the lgdhip C-ABI, or,
// for a cell without one, the cell's own update() compiled for the GPU by
// the generic family (libgeodecomp/hip/genericfamily.h, needs hipcc); the
// host only holds the Grid that Writers and Steerers see.  A
// struct-of-arrays cell (HasSoA + HasFixedCoordsOnlyUpdate + HasUpdateLineX,
// registered with LIBFLATARRAY_REGISTER_SOA) runs its static updateLineX
// over device SoA grids (libgeodecomp/hip/soafamily.h), the path the
// reference's CUDASimulator leaves empty (cudasimulator.h:592-594).
// Differences from CUDASimulator, all in the user's favour:
//  * steerer edits are uploaded before the next sweep (the reference leaves
//    them on the host, cudasimulator.h:450-465);
//  * between two plugin callbacks run() advances all steps in one call into
//    the family (temporal blocking / multi-step launches), instead of one
//    launch per nano-step;
//  * the device state is downloaded only when a plugin looks at it, and for
//    Writers only the x-rows they read (TrackingGrid's on-demand reads over
//    the family's downloadRegion; the reference copies the whole grid back
//    for every writer call, cudasimulator.h:490-497);
//  * after steerers, only the cells they changed are uploaded (the host grid
//    records every write through GridBase, see storage/trackinggrid.h),
//    when the family supports region uploads and the change is small; a
//    changed edge cell or a large change uploads the whole grid.
#ifndef LIBGEODECOMP_PARALLELIZATION_HIPSIMULATOR_H
#define LIBGEODECOMP_PARALLELIZATION_HIPSIMULATOR_H

#include <libgeodecomp/hip/kernelfamilies.h>
#include <libgeodecomp/parallelization/monolithicsimulator.h>
#include <libgeodecomp/storage/grid.h>
#include <libgeodecomp/storage/trackinggrid.h>

#include <memory>
#include <type_traits>

namespace LibGeoDecomp {

namespace HipKernels {

// defined in libgeodecomp/hip/genericfamily.h and libgeodecomp/hip/soafamily.h
// (hipcc translation units)
template<typename CELL>
class Generic;
template<typename CELL>
class SoA;

template<typename T, typename = void>
class IsComplete : public std::false_type {};
template<typename T>
class IsComplete<T, decltype(void(sizeof(T)))> : public std::true_type {};

// does FAMILY offer downloadRegion(&grid, region)?
template<typename FAMILY, typename GRID, typename REGION, typename = void>
class HasDownloadRegion : public std::false_type {};
template<typename FAMILY, typename GRID, typename REGION>
class HasDownloadRegion<FAMILY, GRID, REGION,
                        decltype(std::declval<FAMILY&>().downloadRegion(std::declval<GRID *>(),
                                                                        std::declval<const REGION&>()))>
    : public std::true_type {};

// does FAMILY offer uploadRegion(grid, region)?
template<typename FAMILY, typename GRID, typename REGION, typename = void>
class HasUploadRegion : public std::false_type {};
template<typename FAMILY, typename GRID, typename REGION>
class HasUploadRegion<FAMILY, GRID, REGION,
                      decltype(std::declval<FAMILY&>().uploadRegion(std::declval<const GRID&>(),
                                                                    std::declval<const REGION&>()))>
    : public std::true_type {};

}

template<typename CELL_TYPE>
class HipSimulator : public MonolithicSimulator<CELL_TYPE>
{
public:
    typedef typename MonolithicSimulator<CELL_TYPE>::Topology Topology;
    typedef typename APITraits::SelectHipKernel<CELL_TYPE>::Value RegisteredFamily;
    // a registered hand-written family; else the cell's own code: its static
    // updateLineX over struct-of-arrays grids for SoA line cells, its
    // update() (or array-of-structs updateLineX) for every other cell -- the
    // choice of the reference's UpdateFunctor selector
    // (storage/updatefunctor.h:96-174)
    typedef typename std::conditional<APITraits::SelectSoAStreakUpdate<CELL_TYPE>::VALUE, HipKernels::SoA<CELL_TYPE>,
                                      HipKernels::Generic<CELL_TYPE> >::type OwnCodeFamily;
    typedef typename std::conditional<std::is_void<RegisteredFamily>::value, OwnCodeFamily,
                                      RegisteredFamily>::type Family;
    typedef TrackingGrid<CELL_TYPE, Topology> GridType;
    typedef typename Steerer<CELL_TYPE>::SteererFeedback SteererFeedback;
    static const int DIM = Topology::DIM;
    static const unsigned NANO_STEPS = APITraits::SelectNanoSteps<CELL_TYPE>::VALUE;

    static_assert(HipKernels::IsComplete<Family>::value,
                  "no hand-written HIP kernel family is registered for this cell: include "
                  "<libgeodecomp/hip/genericfamily.h> and compile with hipcc to run the cell's own update() or "
                  "updateLineX() on the GPU, or name a family with APITraits::HasHipKernel<...> / "
                  "LIBGEODECOMP_REGISTER_HIP_KERNEL");

    using MonolithicSimulator<CELL_TYPE>::chronometer;
    using MonolithicSimulator<CELL_TYPE>::initializer;
    using MonolithicSimulator<CELL_TYPE>::steerers;
    using MonolithicSimulator<CELL_TYPE>::stepNum;
    using MonolithicSimulator<CELL_TYPE>::writers;
    using MonolithicSimulator<CELL_TYPE>::gridDim;

    // blockSize keeps the reference constructor's shape
    // (cudasimulator.h:314-316); a zero component leaves the family's
    // tuned default for that knob.  stepsPerLaunch > 0 sets the temporal
    // blocking depth of families that have one.
    explicit HipSimulator(Initializer<CELL_TYPE> *initializer, Coord<3> blockSize = Coord<3>(0, 0, 0),
                          int stepsPerLaunch = 0, int variant = 0) :
        MonolithicSimulator<CELL_TYPE>(initializer),
        hostGrid(initializer->gridBox()),
        hostCurrent(true),
        deviceCurrent(false)
    {
        HipKernels::Tuning tuning;
        if (blockSize.x() || blockSize.y() || blockSize.z() || stepsPerLaunch || variant) {
            tuning.useTile = true;
            tuning.tile.bx = blockSize.x();
            tuning.tile.by = blockSize.y();
            tuning.tile.bz = blockSize.z();
            tuning.tile.steps_per_launch = stepsPerLaunch;
            tuning.tile.variant = variant;
        }
        device.reset(new Family(initializer->gridBox().dimensions, tuning));
        stepNum = initializer->startStep();
        initializer->grid(&hostGrid);
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
        advance(1);
        ++stepNum;
        handleOutput(stepNum == initializer->maxSteps() ? WRITER_ALL_DONE : WRITER_STEP_FINISHED);
    }

    virtual void run()
    {
        initializer->grid(&hostGrid);
        hostGrid.markCurrent();
        hostCurrent = true;
        writerView = false;
        deviceCurrent = false;
        stepNum = initializer->startStep();
        setIORegions();
        SteererFeedback feedback;
        handleInput(STEERER_INITIALIZED, &feedback);
        handleOutput(WRITER_INITIALIZED);

        while (stepNum < initializer->maxSteps()) {
            if (feedback.simulationEnded()) {
                break;
            }
            const unsigned quiet = quietSteps();
            if (quiet <= 1) {
                step(&feedback);
                continue;
            }
            // no steerer wakes up before, and no writer after, any of the
            // next `quiet` steps but the last one: run them as one batch
            TimeTotal t(&chronometer);
            advance(quiet);
            stepNum += quiet;
            handleOutput(stepNum == initializer->maxSteps() ? WRITER_ALL_DONE : WRITER_STEP_FINISHED);
        }
        handleInput(STEERER_ALL_DONE, &feedback);
    }

    virtual const typename Simulator<CELL_TYPE>::GridType *getGrid()
    {
        syncHost();
        return &hostGrid;
    }

    // number of device sweeps issued so far (for tests and tracing)
    unsigned long long sweepsIssued() const { return sweeps; }

    // the kernel family holding the device state (for tests and tuning)
    const Family& deviceFamily() const { return *device; }

    // host -> device transfers so far: whole-grid uploads, region uploads
    // after steerers, and the cells those region uploads carried
    unsigned long long fullUploads() const { return nFull; }
    unsigned long long regionUploads() const { return nRegion; }
    unsigned long long regionCellsUploaded() const { return regionCells; }
    // device -> host transfers so far: whole-grid downloads and the x-rows
    // Writers brought in one by one
    unsigned long long fullDownloads() const { return nFullDown; }
    std::size_t rowsDownloaded() const { return hostGrid.rowsFetchedOnDemand(); }

    // a region upload is used when it carries at most this share of the grid
    static constexpr double REGION_UPLOAD_MAX_SHARE = 0.125;

protected:
    GridType hostGrid;
    std::unique_ptr<Family> device;
    bool hostCurrent;
    bool deviceCurrent;
    unsigned long long sweeps = 0;
    bool writerView = false;     // host rows stale, fetched as Writers read them
    Region<DIM> pendingRegion;   // host edits not yet on the device
    unsigned long long nFull = 0, nRegion = 0, regionCells = 0, nFullDown = 0;

    void syncHost()
    {
        if (!hostCurrent) {
            device->download(&hostGrid);
            ++nFullDown;
            hostGrid.markCurrent();
            hostCurrent = true;
            writerView = false;
        }
    }

    // before Writers look at a stale host grid: rows come in as they read
    // them (whole grid at once for a family without region downloads)
    void prepareWriterView()
    {
        if (hostCurrent || writerView) {
            return;
        }
        prepareWriterView(HipKernels::HasDownloadRegion<Family, Grid<CELL_TYPE, Topology>, Region<DIM> >());
    }

    void prepareWriterView(std::true_type)
    {
        hostGrid.markStale([this](const Region<DIM>& rows) { device->downloadRegion(&hostGrid, rows); },
                           [this]() {
                               device->download(&hostGrid);
                               ++nFullDown;
                           });
        writerView = true;
    }

    void prepareWriterView(std::false_type)
    {
        syncHost();
    }

    // after the Writers: every row fetched means the host copy is whole
    void finishWriterView()
    {
        if (writerView && !hostGrid.hasStaleRows()) {
            hostCurrent = true;
            writerView = false;
        }
    }

    void syncDevice()
    {
        if (!deviceCurrent) {
            device->upload(hostGrid);
            deviceCurrent = true;
            pendingRegion = Region<DIM>();
            ++nFull;
            return;
        }
        if (!pendingRegion.empty()) {
            uploadPending(HipKernels::HasUploadRegion<Family, Grid<CELL_TYPE, Topology>, Region<DIM> >());
        }
    }

    void uploadPending(std::true_type)
    {
        device->uploadRegion(hostGrid, pendingRegion);
        ++nRegion;
        regionCells += pendingRegion.size();
        pendingRegion = Region<DIM>();
    }

    void uploadPending(std::false_type)
    {
        device->upload(hostGrid);
        ++nFull;
        pendingRegion = Region<DIM>();
    }

    // after steerers: remember what they changed for the next upload
    void noteSteererEdits()
    {
        if (!deviceCurrent) {
            return;   // a whole-grid upload is due anyway
        }
        const Region<DIM>& changed = hostGrid.changedRegion();
        if (hostGrid.edgeWasChanged()) {
            deviceCurrent = false;
            return;
        }
        const std::vector<Streak<DIM> > streaks = changed.streaks();
        for (std::size_t i = 0; i < streaks.size(); ++i) {
            pendingRegion << streaks[i];
        }
        const bool region = HipKernels::HasUploadRegion<Family, Grid<CELL_TYPE, Topology>, Region<DIM> >::value;
        if (!region || (double)pendingRegion.size() > REGION_UPLOAD_MAX_SHARE * (double)hostGrid.size()) {
            deviceCurrent = false;
        }
    }

    void advance(unsigned steps)
    {
        TimeCompute t(&chronometer);
        syncDevice();
        device->run(steps * NANO_STEPS);
        device->sync();
        sweeps += (unsigned long long)steps * NANO_STEPS;
        hostCurrent = false;
        writerView = false;
    }

    // how many steps from stepNum on can run without a plugin callback in
    // between (at least 1)
    unsigned quietSteps() const
    {
        const unsigned last = initializer->maxSteps();
        unsigned k = 0;
        for (unsigned s = stepNum; s < last; ++s) {
            if (s != stepNum) {
                for (std::size_t i = 0; i < steerers.size(); ++i) {
                    if (s % steerers[i]->getPeriod() == 0) {
                        return k;
                    }
                }
            } else {
                for (std::size_t i = 0; i < steerers.size(); ++i) {
                    if (s % steerers[i]->getPeriod() == 0) {
                        return 0;
                    }
                }
            }
            ++k;
            const unsigned after = s + 1;
            for (std::size_t i = 0; i < writers.size(); ++i) {
                if (after % writers[i]->getPeriod() == 0 || after == last) {
                    return k;
                }
            }
        }
        return k;
    }

    void handleOutput(WriterEvent event)
    {
        TimeOutput t(&chronometer);
        for (std::size_t i = 0; i < writers.size(); ++i) {
            if ((event != WRITER_STEP_FINISHED) || ((stepNum % writers[i]->getPeriod()) == 0)) {
                prepareWriterView();
                writers[i]->stepFinished(hostGrid, stepNum, event);
            }
        }
        finishWriterView();
    }

    void handleInput(SteererEvent event, SteererFeedback *feedback)
    {
        TimeInput t(&chronometer);
        Region<DIM> simArea;
        simArea << hostGrid.boundingBox();
        bool any = false;
        for (std::size_t i = 0; i < steerers.size(); ++i) {
            if ((event != STEERER_NEXT_STEP) || (stepNum % steerers[i]->getPeriod() == 0)) {
                syncHost();
                if (!any) {
                    hostGrid.startTracking();
                    any = true;
                }
                steerers[i]->nextStep(&hostGrid, simArea, gridDim, stepNum, event, 0, true, feedback);
            }
        }
        if (any) {
            // the steerers may have changed cells: upload them before the
            // next sweep
            hostGrid.stopTracking();
            noteSteererEdits();
        }
    }

    void setIORegions()
    {
        Region<DIM> simArea;
        simArea << hostGrid.boundingBox();
        for (std::size_t i = 0; i < steerers.size(); ++i) {
            steerers[i]->setRegion(simArea);
        }
    }
};

}

#endif
