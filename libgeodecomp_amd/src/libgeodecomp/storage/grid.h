// Grid<CELL, TOPOLOGY>: host array-of-structs grid over a CoordBox with a
// constant edge cell returned for accesses outside a non-wrapping axis
// (src/libgeodecomp/storage/grid.h; topology lookup as in
// geometry/topologies.h:186-199).
#ifndef LIBGEODECOMP_STORAGE_GRID_H
#define LIBGEODECOMP_STORAGE_GRID_H

#include <libgeodecomp/geometry/topologies.h>
#include <libgeodecomp/storage/gridbase.h>

#include <stdexcept>
#include <vector>

namespace LibGeoDecomp {

template<typename CELL, typename TOPOLOGY = Topologies::Cube<2>::Topology>
class Grid : public GridBase<CELL, TOPOLOGY::DIM>
{
public:
    typedef TOPOLOGY Topology;
    static const int DIM = TOPOLOGY::DIM;

    explicit Grid(const CoordBox<DIM>& box = CoordBox<DIM>(), const CELL& defaultCell = CELL(),
                  const CELL& edgeCell = CELL()) :
        box(box),
        cells(box.size(), defaultCell),
        edgeCell(edgeCell)
    {}

    explicit Grid(const Coord<DIM>& dim, const CELL& defaultCell = CELL(), const CELL& edgeCell = CELL()) :
        box(Coord<DIM>(), dim),
        cells(box.size(), defaultCell),
        edgeCell(edgeCell)
    {}

    void resize(const CoordBox<DIM>& newBox)
    {
        box = newBox;
        cells.assign(box.size(), CELL());
    }

    // neighbourhood access: logical coordinate, topology-normalised
    const CELL& operator[](const Coord<DIM>& c) const
    {
        const Coord<DIM> rel = c - box.origin;
        if (Topology::isOutOfBounds(rel, box.dimensions)) {
            return edgeCell;
        }
        return cells[Topology::normalize(rel, box.dimensions).toIndex(box.dimensions)];
    }

    CELL& operator[](const Coord<DIM>& c)
    {
        const Coord<DIM> rel = c - box.origin;
        if (Topology::isOutOfBounds(rel, box.dimensions)) {
            return edgeCell;
        }
        return cells[Topology::normalize(rel, box.dimensions).toIndex(box.dimensions)];
    }

    void set(const Coord<DIM>& c, const CELL& cell) { (*this)[c] = cell; }

    void set(const Streak<DIM>& s, const CELL *src)
    {
        Coord<DIM> c = s.origin;
        for (; c.x() < s.endX; ++c.x()) {
            (*this)[c] = *src++;
        }
    }

    CELL get(const Coord<DIM>& c) const { return (*this)[c]; }

    void get(const Streak<DIM>& s, CELL *dst) const
    {
        Coord<DIM> c = s.origin;
        for (; c.x() < s.endX; ++c.x()) {
            *dst++ = (*this)[c];
        }
    }

    void setEdge(const CELL& e) { edgeCell = e; }
    const CELL& getEdge() const { return edgeCell; }
    CoordBox<DIM> boundingBox() const { return box; }

    CELL *data() { return cells.empty() ? 0 : &cells[0]; }
    const CELL *data() const { return cells.empty() ? 0 : &cells[0]; }
    std::size_t size() const { return cells.size(); }

private:
    CoordBox<DIM> box;
    std::vector<CELL> cells;
    CELL edgeCell;
};

}

#endif
