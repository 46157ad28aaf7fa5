// TrackingGrid<CELL, TOPOLOGY>: a host Grid that records which cells are
// written through the GridBase interface (set() of a coordinate or a
// streak, setEdge()) while tracking is on.  Steerers can only change a grid
// through that interface (Steerer::nextStep receives a GridBase*,
// src/libgeodecomp/io/steerer.h), so HipSimulator hands them this grid and
// uploads just the changed streaks afterwards instead of the whole grid.
// Coordinates are recorded topology-normalised; a write outside an
// unwrapped axis lands on the edge cell and is recorded as an edge change.
#ifndef LIBGEODECOMP_STORAGE_TRACKINGGRID_H
#define LIBGEODECOMP_STORAGE_TRACKINGGRID_H

#include <libgeodecomp/geometry/region.h>
#include <libgeodecomp/storage/grid.h>

namespace LibGeoDecomp {

template<typename CELL, typename TOPOLOGY>
class TrackingGrid : public Grid<CELL, TOPOLOGY>
{
public:
    typedef Grid<CELL, TOPOLOGY> Base;
    static const int DIM = TOPOLOGY::DIM;

    explicit TrackingGrid(const CoordBox<DIM>& box = CoordBox<DIM>()) :
        Base(box),
        tracking(false),
        edgeChanged(false)
    {}

    void startTracking()
    {
        tracking = true;
        edgeChanged = false;
        changed = Region<DIM>();
    }

    void stopTracking() { tracking = false; }

    const Region<DIM>& changedRegion() const { return changed; }
    bool edgeWasChanged() const { return edgeChanged; }

    virtual void set(const Coord<DIM>& c, const CELL& cell)
    {
        note(c);
        Base::set(c, cell);
    }

    virtual void set(const Streak<DIM>& s, const CELL *src)
    {
        if (tracking) {
            const CoordBox<DIM> box = this->boundingBox();
            Coord<DIM> last = s.origin;
            last.x() = s.endX - 1;
            if (s.endX > s.origin.x() && box.inBounds(s.origin) && box.inBounds(last)) {
                changed << s;
            } else {
                for (Coord<DIM> c = s.origin; c.x() < s.endX; ++c.x()) {
                    note(c);
                }
            }
        }
        Base::set(s, src);
    }

    virtual void setEdge(const CELL& e)
    {
        if (tracking) {
            edgeChanged = true;
        }
        Base::setEdge(e);
    }

private:
    bool tracking;
    bool edgeChanged;
    Region<DIM> changed;

    void note(const Coord<DIM>& c)
    {
        if (!tracking) {
            return;
        }
        const CoordBox<DIM> box = this->boundingBox();
        const Coord<DIM> rel = c - box.origin;
        if (TOPOLOGY::isOutOfBounds(rel, box.dimensions)) {
            edgeChanged = true;
            return;
        }
        changed << (TOPOLOGY::normalize(rel, box.dimensions) + box.origin);
    }
};

}

#endif
