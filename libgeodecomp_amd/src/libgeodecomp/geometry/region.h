// Region<DIM>: a set of cells kept as disjoint streaks.  This is the small
// subset of src/libgeodecomp/geometry/region.h that the monolithic
// simulators hand to Steerers (validRegion) and Writers: insertion of boxes
// and streaks, bounding box, size and coordinate iteration.
#ifndef LIBGEODECOMP_GEOMETRY_REGION_H
#define LIBGEODECOMP_GEOMETRY_REGION_H

#include <libgeodecomp/geometry/coordbox.h>
#include <libgeodecomp/geometry/streak.h>

#include <algorithm>
#include <map>
#include <vector>

namespace LibGeoDecomp {

template<int DIM>
class Region
{
public:
    typedef std::vector<Coord<DIM> > CoordVector;

    Region& operator<<(const Streak<DIM>& s)
    {
        if (s.endX <= s.origin.x()) {
            return *this;
        }
        Coord<DIM> key = s.origin;
        key.x() = 0;
        std::vector<std::pair<int, int> >& row = rows[key];
        row.push_back(std::make_pair(s.origin.x(), s.endX));
        normalizeRow(&row);
        return *this;
    }

    Region& operator<<(const CoordBox<DIM>& box)
    {
        if (box.size() == 0) {
            return *this;
        }
        CoordBox<DIM> rowsBox = box;
        rowsBox.dimensions.x() = 1;
        for (typename CoordBox<DIM>::Iterator i = rowsBox.begin(); i != rowsBox.end(); ++i) {
            *this << Streak<DIM>(*i, i->x() + box.dimensions.x());
        }
        return *this;
    }

    Region& operator<<(const Coord<DIM>& c)
    {
        return *this << Streak<DIM>(c, c.x() + 1);
    }

    std::size_t size() const
    {
        std::size_t n = 0;
        for (typename RowMap::const_iterator r = rows.begin(); r != rows.end(); ++r) {
            for (std::size_t k = 0; k < r->second.size(); ++k) {
                n += r->second[k].second - r->second[k].first;
            }
        }
        return n;
    }

    bool empty() const { return rows.empty(); }

    CoordBox<DIM> boundingBox() const
    {
        if (rows.empty()) {
            return CoordBox<DIM>();
        }
        Coord<DIM> lo = Coord<DIM>::diagonal(1 << 30);
        Coord<DIM> hi = Coord<DIM>::diagonal(-(1 << 30));
        for (typename RowMap::const_iterator r = rows.begin(); r != rows.end(); ++r) {
            for (int d = 1; d < DIM; ++d) {
                lo[d] = std::min(lo[d], r->first[d]);
                hi[d] = std::max(hi[d], r->first[d] + 1);
            }
            lo[0] = std::min(lo[0], r->second.front().first);
            hi[0] = std::max(hi[0], r->second.back().second);
        }
        return CoordBox<DIM>(lo, hi - lo);
    }

    bool count(const Coord<DIM>& c) const
    {
        Coord<DIM> key = c;
        key.x() = 0;
        typename RowMap::const_iterator r = rows.find(key);
        if (r == rows.end()) {
            return false;
        }
        for (std::size_t k = 0; k < r->second.size(); ++k) {
            if (c.x() >= r->second[k].first && c.x() < r->second[k].second) {
                return true;
            }
        }
        return false;
    }

    std::vector<Streak<DIM> > streaks() const
    {
        std::vector<Streak<DIM> > out;
        for (typename RowMap::const_iterator r = rows.begin(); r != rows.end(); ++r) {
            for (std::size_t k = 0; k < r->second.size(); ++k) {
                Coord<DIM> o = r->first;
                o.x() = r->second[k].first;
                out.push_back(Streak<DIM>(o, r->second[k].second));
            }
        }
        return out;
    }

    CoordVector toVector() const
    {
        CoordVector out;
        std::vector<Streak<DIM> > s = streaks();
        for (std::size_t k = 0; k < s.size(); ++k) {
            for (int x = s[k].origin.x(); x < s[k].endX; ++x) {
                Coord<DIM> c = s[k].origin;
                c.x() = x;
                out.push_back(c);
            }
        }
        return out;
    }

    bool operator==(const Region& o) const { return rows == o.rows; }

private:
    struct RowLess
    {
        bool operator()(const Coord<DIM>& a, const Coord<DIM>& b) const { return a < b; }
    };
    typedef std::map<Coord<DIM>, std::vector<std::pair<int, int> >, RowLess> RowMap;
    RowMap rows;

    static void normalizeRow(std::vector<std::pair<int, int> > *row)
    {
        std::sort(row->begin(), row->end());
        std::vector<std::pair<int, int> > merged;
        for (std::size_t k = 0; k < row->size(); ++k) {
            if (!merged.empty() && (*row)[k].first <= merged.back().second) {
                merged.back().second = std::max(merged.back().second, (*row)[k].second);
            } else {
                merged.push_back((*row)[k]);
            }
        }
        row->swap(merged);
    }
};

}

#endif
