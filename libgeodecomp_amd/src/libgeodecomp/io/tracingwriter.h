// TracingWriter<CELL>: progress and performance log of a run, the
// reference's plugin of the same name (src/libgeodecomp/io/tracingwriter.h:
// 27-167): on WRITER_INITIALIZED it notes the start time; every `period`
// steps it prints the step, elapsed time, remaining time and ETA, the update
// rate in GLUPS = steps * NANO_STEPS * cells / elapsed / 1e9 and the
// effective memory bandwidth GLUPS * 2 * sizeof(CELL) GB/s
// (tracingwriter.h:124-125: every cell read once and written once per
// update); on WRITER_ALL_DONE the total time.
//
// It reads no cell of the grid, only its dimensions, so under HipSimulator a
// TracingWriter costs no device-to-host transfer at all (the simulator's host
// grid fetches rows only when a writer reads them).  Rates count the steps
// since WRITER_INITIALIZED (the reference divides by the absolute step
// number, which overstates the rate of a run resumed at a later step).
#ifndef LIBGEODECOMP_IO_TRACINGWRITER_H
#define LIBGEODECOMP_IO_TRACINGWRITER_H

#include <libgeodecomp/io/writer.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>

namespace LibGeoDecomp {

template<typename CELL_TYPE>
class TracingWriter : public Writer<CELL_TYPE>
{
public:
    typedef typename Writer<CELL_TYPE>::GridType GridType;
    using Writer<CELL_TYPE>::NANO_STEPS;
    static const int DIM = Writer<CELL_TYPE>::DIM;
    // the reference's rank selector (only one rank exists here)
    static const int OUTPUT_ON_ALL_RANKS = -1;

    explicit TracingWriter(const unsigned period = 1, const unsigned maxSteps = 1,
                           int outputRank = OUTPUT_ON_ALL_RANKS, std::ostream& stream = std::cerr) :
        Writer<CELL_TYPE>("", period),
        outputRank(outputRank),
        stream(stream),
        startTime(0),
        startStep(0),
        maxSteps(maxSteps)
    {}

    Writer<CELL_TYPE> *clone() const { return new TracingWriter(*this); }

    void stepFinished(const GridType& grid, unsigned step, WriterEvent event)
    {
        if (outputRank != OUTPUT_ON_ALL_RANKS && outputRank != 0) {
            return;
        }
        switch (event) {
        case WRITER_INITIALIZED:
            startTime = now();
            startStep = step;
            stream << "TracingWriter::initialized()\n";
            printTime();
            break;
        case WRITER_STEP_FINISHED:
            if (step % this->period == 0) {
                progress(step, grid.dimensions());
            }
            break;
        case WRITER_ALL_DONE:
            stream << "TracingWriter::allDone()\n"
                   << "  total time: " << renderDuration(now() - startTime) << "\n";
            printTime();
            break;
        default:
            throw std::invalid_argument("unknown writer event");
        }
    }

    // seconds since an arbitrary fixed point, microsecond resolution or better
    static double now()
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }

    // hh:mm:ss with microseconds when the duration has a fraction
    static std::string renderDuration(double seconds)
    {
        if (!(seconds >= 0)) {
            seconds = 0;
        }
        const long whole = (long)std::floor(seconds);
        const long micro = (long)((seconds - (double)whole) * 1e6);
        char buf[64];
        if (micro > 0) {
            std::snprintf(buf, sizeof(buf), "%02ld:%02ld:%02ld.%06ld", whole / 3600, whole / 60 % 60, whole % 60,
                          micro);
        } else {
            std::snprintf(buf, sizeof(buf), "%02ld:%02ld:%02ld", whole / 3600, whole / 60 % 60, whole % 60);
        }
        return buf;
    }

    // the rates of `steps` steps over `cells` cells in `seconds`
    static double glups(double steps, double cells, double seconds)
    {
        return seconds > 0 ? steps * NANO_STEPS * cells / seconds * 1e-9 : 0.0;
    }
    static double bandwidthGBs(double glupsValue) { return glupsValue * 2.0 * sizeof(CELL_TYPE); }

private:
    int outputRank;
    std::ostream& stream;
    double startTime;
    unsigned startStep;
    unsigned maxSteps;

    void progress(unsigned step, const Coord<DIM>& dim)
    {
        const double t = now();
        const double elapsed = t - startTime;
        const double done = (double)step - (double)startStep;
        const double left = (double)maxSteps > (double)step ? (double)maxSteps - (double)step : 0.0;
        const double remaining = done > 0 ? elapsed * left / done : 0.0;
        const double rate = glups(done, (double)dim.prod(), elapsed);
        stream << "TracingWriter::stepFinished()\n"
               << "  step: " << step << " of " << maxSteps << "\n"
               << "  elapsed: " << elapsed << "\n"
               << "  remaining: " << renderDuration(remaining) << "\n"
               << "  ETA:  " << renderDuration(t + remaining) << "\n"
               << "  speed: " << rate << " GLUPS\n"
               << "  effective memory bandwidth " << bandwidthGBs(rate) << " GB/s\n";
        printTime();
    }

    void printTime()
    {
        stream << "  time: " << renderDuration(now()) << "\n";
        stream.flush();
    }
};

}

#endif
