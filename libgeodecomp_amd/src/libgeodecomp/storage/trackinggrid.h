// TrackingGrid<CELL, TOPOLOGY>: a host Grid that records which cells are
// written through the GridBase interface (set() of a coordinate or a
// streak, setEdge()) while tracking is on.  Steerers can only change a grid
// through that interface (Steerer::nextStep receives a GridBase*,
// src/libgeodecomp/io/steerer.h), so HipSimulator hands them this grid and
// uploads just the changed streaks afterwards instead of the whole grid.
// Coordinates are recorded topology-normalised; a write outside an
// unwrapped axis lands on the edge cell and is recorded as an edge change.
//
// It also reads on demand: after markStale() every x-row (one (y, z) line
// of the box) is stale, and the first read of a cell in a stale row through
// the GridBase interface (get() of a coordinate or a streak) first brings
// that row in through the fetch callback.  HipSimulator arms this before
// its Writers run, so a writer downloads only the rows it reads (none for a
// writer that only looks at the dimensions, such as TracingWriter) instead
// of the whole grid (cudasimulator.h:490-497 copies the whole grid back).
// Once more than a share of the rows came one by one, the rest come in one
// whole-grid transfer.  Reads of the edge cell never fetch: the edge is a
// host-side constant.
#ifndef LIBGEODECOMP_STORAGE_TRACKINGGRID_H
#define LIBGEODECOMP_STORAGE_TRACKINGGRID_H

#include <libgeodecomp/geometry/region.h>
#include <libgeodecomp/storage/grid.h>

#include <algorithm>
#include <functional>
#include <vector>

namespace LibGeoDecomp {

template<typename CELL, typename TOPOLOGY>
class TrackingGrid : public Grid<CELL, TOPOLOGY>
{
public:
    typedef Grid<CELL, TOPOLOGY> Base;
    static const int DIM = TOPOLOGY::DIM;

    // fetchRows(region): bring the region's cells (whole x-rows) into the
    // grid; fetchAll(): bring every cell in
    typedef std::function<void(const Region<DIM>&)> FetchRows;
    typedef std::function<void()> FetchAll;

    explicit TrackingGrid(const CoordBox<DIM>& box = CoordBox<DIM>()) :
        Base(box),
        tracking(false),
        edgeChanged(false),
        staleRows(0),
        rowsFetched(0),
        rowsThisTime(0)
    {}

    // every row stale from now on, fetched on first read
    void markStale(const FetchRows& rows, const FetchAll& all)
    {
        fetchRows = rows;
        fetchAll = all;
        stale.assign(rowCount(), 1);
        staleRows = stale.size();
        rowsThisTime = 0;
    }

    // the host copy is whole again (after a whole-grid transfer or an edit)
    void markCurrent()
    {
        stale.clear();
        staleRows = 0;
    }

    bool hasStaleRows() const { return staleRows != 0; }
    // rows brought in one by one so far (for tests and tuning)
    std::size_t rowsFetchedOnDemand() const { return rowsFetched; }

    // rows fetched one by one before the rest come in one transfer
    std::size_t rowLimit() const { return std::max<std::size_t>(16, rowCount() / 32); }

    virtual CELL get(const Coord<DIM>& c) const
    {
        touch(c);
        return Base::get(c);
    }

    virtual void get(const Streak<DIM>& s, CELL *dst) const
    {
        // every cell of a streak lies in one (y, z) row or on the edge
        if (s.endX > s.origin.x()) {
            touch(s.origin);
        }
        Base::get(s, dst);
    }

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
    FetchRows fetchRows;
    FetchAll fetchAll;
    // one flag per x-row while reads are on demand (cache state of a
    // logically const grid, hence mutable)
    mutable std::vector<unsigned char> stale;
    mutable std::size_t staleRows;
    mutable std::size_t rowsFetched;
    mutable std::size_t rowsThisTime;   // since the last markStale()

    std::size_t rowCount() const
    {
        const Coord<DIM> d = this->boundingBox().dimensions;
        std::size_t rows = 1;
        for (int i = 1; i < DIM; ++i) {
            rows *= (std::size_t)d[i];
        }
        return d.prod() == 0 ? 0 : rows;
    }

    void touch(const Coord<DIM>& c) const
    {
        if (!staleRows) {
            return;
        }
        const CoordBox<DIM> box = this->boundingBox();
        const Coord<DIM> rel = c - box.origin;
        if (TOPOLOGY::isOutOfBounds(rel, box.dimensions)) {
            return;
        }
        const Coord<DIM> n = TOPOLOGY::normalize(rel, box.dimensions);
        std::size_t row = 0;
        for (int i = DIM - 1; i >= 1; --i) {
            row = row * (std::size_t)box.dimensions[i] + (std::size_t)n[i];
        }
        if (!stale[row]) {
            return;
        }
        if (rowsThisTime >= rowLimit()) {
            fetchAll();
            stale.clear();
            staleRows = 0;
            return;
        }
        Coord<DIM> origin = n + box.origin;
        origin.x() = box.origin.x();
        Region<DIM> r;
        r << Streak<DIM>(origin, box.origin.x() + box.dimensions.x());
        fetchRows(r);
        stale[row] = 0;
        --staleRows;
        ++rowsFetched;
        ++rowsThisTime;
    }

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
