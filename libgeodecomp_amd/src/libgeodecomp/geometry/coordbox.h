// CoordBox<DIM>: axis-aligned box (origin, dimensions) with an x-fastest
// iterator and a row (streak) iterator (src/libgeodecomp/geometry/coordbox.h).
#ifndef LIBGEODECOMP_GEOMETRY_COORDBOX_H
#define LIBGEODECOMP_GEOMETRY_COORDBOX_H

#include <libgeodecomp/geometry/coord.h>
#include <libgeodecomp/geometry/streak.h>

namespace LibGeoDecomp {

template<int DIM>
class CoordBox
{
public:
    Coord<DIM> origin;
    Coord<DIM> dimensions;

    explicit CoordBox(const Coord<DIM>& origin = Coord<DIM>(), const Coord<DIM>& dimensions = Coord<DIM>()) :
        origin(origin),
        dimensions(dimensions)
    {}

    bool inBounds(const Coord<DIM>& c) const
    {
        for (int d = 0; d < DIM; ++d) {
            if (c[d] < origin[d] || c[d] >= origin[d] + dimensions[d]) {
                return false;
            }
        }
        return true;
    }

    std::size_t size() const { return dimensions.prod() < 0 ? 0 : static_cast<std::size_t>(dimensions.prod()); }

    bool operator==(const CoordBox& o) const { return origin == o.origin && dimensions == o.dimensions; }
    bool operator!=(const CoordBox& o) const { return !(*this == o); }

    class Iterator
    {
    public:
        Iterator(const Coord<DIM>& origin, const Coord<DIM>& start, const Coord<DIM>& dimensions) :
            origin(origin), cursor(start), end(origin + dimensions)
        {}

        const Coord<DIM>& operator*() const { return cursor; }
        const Coord<DIM>* operator->() const { return &cursor; }
        bool operator==(const Iterator& o) const { return cursor == o.cursor; }
        bool operator!=(const Iterator& o) const { return cursor != o.cursor; }

        Iterator& operator++()
        {
            for (int d = 0; d < DIM; ++d) {
                if (++cursor[d] < end[d]) {
                    return *this;
                }
                if (d < DIM - 1) {
                    cursor[d] = origin[d];
                }
            }
            return *this;
        }

    private:
        Coord<DIM> origin;
        Coord<DIM> cursor;
        Coord<DIM> end;
    };

    Iterator begin() const
    {
        if (size() == 0) {
            return end();
        }
        return Iterator(origin, origin, dimensions);
    }

    Iterator end() const
    {
        Coord<DIM> e = origin;
        if (size() != 0) {
            e[DIM - 1] = origin[DIM - 1] + dimensions[DIM - 1];
        }
        return Iterator(origin, e, dimensions);
    }

    // the box's rows along x, y fastest then z: one Streak per row
    class StreakIterator
    {
    public:
        StreakIterator(const CoordBox& box, const Coord<DIM>& start) : box(box), cursor(start) {}

        Streak<DIM> operator*() const { return Streak<DIM>(cursor, cursor.x() + box.dimensions.x()); }
        bool operator==(const StreakIterator& o) const { return cursor == o.cursor; }
        bool operator!=(const StreakIterator& o) const { return !(cursor == o.cursor); }

        StreakIterator& operator++()
        {
            for (int d = 1; d < DIM; ++d) {
                if (++cursor[d] < box.origin[d] + box.dimensions[d]) {
                    return *this;
                }
                if (d < DIM - 1) {
                    cursor[d] = box.origin[d];
                }
            }
            if (DIM == 1) {
                cursor[0] = box.origin[0] + box.dimensions[0];
            }
            return *this;
        }

    private:
        CoordBox box;
        Coord<DIM> cursor;
    };

    StreakIterator beginStreak() const { return size() == 0 ? endStreak() : StreakIterator(*this, origin); }

    StreakIterator endStreak() const
    {
        Coord<DIM> e = origin;
        if (size() != 0) {
            if (DIM == 1) {
                e[0] = origin[0] + dimensions[0];
            } else {
                e[DIM - 1] = origin[DIM - 1] + dimensions[DIM - 1];
            }
        }
        return StreakIterator(*this, e);
    }
};

}

#endif
