// Chronometer: accumulates wall time per simulator phase.  Event classes
// and names follow src/libgeodecomp/misc/chronometer.h:142-150
// (TimeTotal ... TimeOutput, "total_time" ... "output_time"); the RAII
// timers add their elapsed time to the chronometer on destruction.
#ifndef LIBGEODECOMP_MISC_CHRONOMETER_H
#define LIBGEODECOMP_MISC_CHRONOMETER_H

#include <chrono>
#include <sstream>
#include <string>

namespace LibGeoDecomp {

class Chronometer
{
public:
    static const int NUM_EVENTS = 9;

    Chronometer() { reset(); }

    void reset()
    {
        for (int i = 0; i < NUM_EVENTS; ++i) {
            totalTimes[i] = 0;
        }
    }

    template<typename EVENT>
    void addTime(double t) { totalTimes[EVENT::ID] += t; }

    template<typename EVENT>
    double interval() const { return totalTimes[EVENT::ID]; }

    template<typename EVENT1, typename EVENT2>
    double ratio() const
    {
        double d = totalTimes[EVENT2::ID];
        return d == 0 ? 0 : totalTimes[EVENT1::ID] / d;
    }

    double *rawTotalTimes() { return totalTimes; }

    Chronometer& operator+=(const Chronometer& o)
    {
        for (int i = 0; i < NUM_EVENTS; ++i) {
            totalTimes[i] += o.totalTimes[i];
        }
        return *this;
    }

    static std::string eventName(int id)
    {
        static const char *names[NUM_EVENTS] = {
            "total_time", "compute_time", "compute_time_inner", "compute_time_ghost",
            "patch_accepters_time", "patch_providers_time", "communication_time", "input_time",
            "output_time"};
        return (id >= 0 && id < NUM_EVENTS) ? names[id] : "unknown";
    }

    std::string report() const
    {
        std::ostringstream s;
        for (int i = 0; i < NUM_EVENTS; ++i) {
            s << eventName(i) << " " << totalTimes[i] << "\n";
        }
        return s.str();
    }

    static double now()
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }

private:
    double totalTimes[NUM_EVENTS];
};

namespace ChronometerHelpers {

template<int EVENT_ID>
class ScopedEvent
{
public:
    static const int ID = EVENT_ID;

    explicit ScopedEvent(Chronometer *c) : chrono(c), start(Chronometer::now()) {}
    ~ScopedEvent() { chrono->rawTotalTimes()[ID] += Chronometer::now() - start; }

private:
    Chronometer *chrono;
    double start;
};

}

#define LIBGEODECOMP_CHRONO_EVENT(NAME, ID)                                           \
    class NAME : public ChronometerHelpers::ScopedEvent<ID>                           \
    {                                                                                 \
    public:                                                                           \
        explicit NAME(Chronometer *c) : ChronometerHelpers::ScopedEvent<ID>(c) {}     \
    };

LIBGEODECOMP_CHRONO_EVENT(TimeTotal, 0)
LIBGEODECOMP_CHRONO_EVENT(TimeCompute, 1)
LIBGEODECOMP_CHRONO_EVENT(TimeComputeInner, 2)
LIBGEODECOMP_CHRONO_EVENT(TimeComputeGhost, 3)
LIBGEODECOMP_CHRONO_EVENT(TimePatchAccepters, 4)
LIBGEODECOMP_CHRONO_EVENT(TimePatchProviders, 5)
LIBGEODECOMP_CHRONO_EVENT(TimeCommunication, 6)
LIBGEODECOMP_CHRONO_EVENT(TimeInput, 7)
LIBGEODECOMP_CHRONO_EVENT(TimeOutput, 8)

#undef LIBGEODECOMP_CHRONO_EVENT

}

#endif
