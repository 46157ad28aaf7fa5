// SoAGrid<CELL, TOPOLOGY>: host grid of a struct-of-arrays cell
// (registered with LIBFLATARRAY_REGISTER_SOA), after
// src/libgeodecomp/storage/soagrid.h:340-794.
//
// Storage: member-major planes over the logical box grown by the edge radii
// of calcEdgeRadii() (soagrid.h:578-584) -- the stencil radius on every
// axis that does not wrap, none on a wrapping axis, whose neighbours are
// reached through FixedNeighborhood's wrap offsets instead.  The padding
// holds the edge cell.  Element (x, y, z) of the logical box is at
//     ((z + rz) * DIM_Y + (y + ry)) * DIM_X + (x + rx)
// and member m of element i at data + DIM_PROD * offset(m) + i * sizeof(m).
//
// SoALayout describes that addressing; the HIP SoA kernel family uses the
// same class with aligned rows (x origin and row pitch rounded up so rows
// start on 128-byte lines) and a plane size rounded up so every member
// plane starts aligned.
#ifndef LIBGEODECOMP_STORAGE_SOAGRID_H
#define LIBGEODECOMP_STORAGE_SOAGRID_H

#include <libflatarray/flat_array.hpp>
#include <libgeodecomp/geometry/topologies.h>
#include <libgeodecomp/misc/apitraits.h>
#include <libgeodecomp/storage/gridbase.h>

#include <vector>

namespace LibGeoDecomp {

class SoALayout
{
public:
    SoALayout() : dimX(1), dimY(1), dimZ(1), prod(1) {}

    // logical extents and edge radii per axis (1 and 0 past the grid's
    // dimensionality); x origin and row pitch rounded up to xAlign
    // elements, the plane size to prodAlign elements
    SoALayout(const Coord<3>& logical, const Coord<3>& radii, long xAlign = 1, long prodAlign = 1) :
        logical(logical),
        radii(radii)
    {
        const long ox = roundUp(radii.x(), xAlign);
        origin = Coord<3>((int)ox, radii.y(), radii.z());
        dimX = roundUp(ox + logical.x() + radii.x(), xAlign);
        dimY = logical.y() + 2L * radii.y();
        dimZ = logical.z() + 2L * radii.z();
        prod = roundUp(dimX * dimY * dimZ, prodAlign);
    }

    long index(long x, long y, long z) const
    {
        return ((z + origin.z()) * dimY + (y + origin.y())) * dimX + (x + origin.x());
    }

    // is the padded element (X, Y, Z) inside the logical box?
    bool inside(long X, long Y, long Z) const
    {
        const long x = X - origin.x(), y = Y - origin.y(), z = Z - origin.z();
        return x >= 0 && x < logical.x() && y >= 0 && y < logical.y() && z >= 0 && z < logical.z();
    }

    Coord<3> logical;
    Coord<3> radii;
    Coord<3> origin;
    long dimX, dimY, dimZ;
    long prod;

private:
    static long roundUp(long v, long a) { return (v + a - 1) / a * a; }
};

namespace SoAGridHelpers {

// the three axes of a DIM-dimensional coordinate (1 past DIM for extents)
template<int DIM>
Coord<3> expand(const Coord<DIM>& c, int fill)
{
    Coord<3> r(fill, fill, fill);
    for (int d = 0; d < DIM; ++d) {
        r[d] = c[d];
    }
    return r;
}

}

template<typename CELL, typename TOPOLOGY = Topologies::Cube<2>::Topology>
class SoAGrid : public GridBase<CELL, TOPOLOGY::DIM>
{
public:
    typedef CELL CellType;
    typedef TOPOLOGY Topology;
    typedef typename APITraits::SelectStencil<CELL>::Value Stencil;
    typedef LibFlatArray::soa_traits<CELL> Traits;
    typedef LibFlatArray::soa_accessor<CELL> Accessor;
    static const int DIM = TOPOLOGY::DIM;

    explicit SoAGrid(const CoordBox<DIM>& box = CoordBox<DIM>(), const CELL& defaultCell = CELL(),
                     const CELL& edgeCell = CELL()) :
        edgeCell(edgeCell)
    {
        resize(box);
        CoordBox<DIM> b = box;
        for (typename CoordBox<DIM>::Iterator i = b.begin(); i != b.end(); ++i) {
            set(*i, defaultCell);
        }
    }

    // edge radius per axis: the stencil radius where the axis does not wrap
    static Coord<3> calcEdgeRadii()
    {
        Coord<3> r;
        for (int d = 0; d < 3; ++d) {
            r[d] = (d < DIM && !Topology::wrapsAxis(d)) ? Stencil::RADIUS : 0;
        }
        return r;
    }

    void resize(const CoordBox<DIM>& newBox)
    {
        box = newBox;
        layout = SoALayout(SoAGridHelpers::expand(box.dimensions, 1), calcEdgeRadii());
        bytes.assign((std::size_t)layout.prod * Traits::BYTES_PER_CELL, 0);
        fillEdge();
    }

    void set(const Coord<DIM>& c, const CELL& cell)
    {
        Coord<DIM> rel = c - box.origin;
        if (Topology::isOutOfBounds(rel, box.dimensions)) {
            setEdge(cell);
            return;
        }
        rel = Topology::normalize(rel, box.dimensions);
        accessorAt(rel) << cell;
    }

    void set(const Streak<DIM>& s, const CELL *cells)
    {
        Coord<DIM> c = s.origin;
        for (; c.x() < s.endX; ++c.x()) {
            set(c, *cells++);
        }
    }

    CELL get(const Coord<DIM>& c) const
    {
        Coord<DIM> rel = c - box.origin;
        if (Topology::isOutOfBounds(rel, box.dimensions)) {
            return edgeCell;
        }
        rel = Topology::normalize(rel, box.dimensions);
        CELL cell;
        accessorAt(rel) >> cell;
        return cell;
    }

    void get(const Streak<DIM>& s, CELL *cells) const
    {
        Coord<DIM> c = s.origin;
        for (; c.x() < s.endX; ++c.x()) {
            *cells++ = get(c);
        }
    }

    // the edge cell is also written into every padding element
    void setEdge(const CELL& cell)
    {
        edgeCell = cell;
        fillEdge();
    }

    const CELL& getEdge() const { return edgeCell; }
    CoordBox<DIM> boundingBox() const { return box; }
    Coord<3> getEdgeRadii() const { return layout.radii; }
    const SoALayout& soaLayout() const { return layout; }

    char *data() { return bytes.data(); }
    const char *data() const { return bytes.data(); }

    // accessor of element `index` of the storage
    Accessor accessor(long index) const
    {
        return Accessor(const_cast<char *>(bytes.data()), index, layout.dimX, layout.dimX * layout.dimY,
                        layout.prod);
    }

private:
    CoordBox<DIM> box;
    CELL edgeCell;
    SoALayout layout;
    std::vector<char> bytes;

    Accessor accessorAt(const Coord<DIM>& rel) const
    {
        const Coord<3> r = SoAGridHelpers::expand(rel, 0);
        return accessor(layout.index(r.x(), r.y(), r.z()));
    }

    void fillEdge()
    {
        for (long Z = 0; Z < layout.dimZ; ++Z) {
            for (long Y = 0; Y < layout.dimY; ++Y) {
                for (long X = 0; X < layout.dimX; ++X) {
                    if (!layout.inside(X, Y, Z)) {
                        Accessor a = accessor((Z * layout.dimY + Y) * layout.dimX + X);
                        a << edgeCell;
                    }
                }
            }
        }
    }
};

}

#endif
