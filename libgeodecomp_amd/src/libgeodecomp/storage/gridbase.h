// GridBase<CELL, DIM>: abstract grid interface handed to Initializers,
// Writers and Steerers (src/libgeodecomp/storage/gridbase.h:71-183):
// single-cell and streak get/set, edge cell, bounding box.
#ifndef LIBGEODECOMP_STORAGE_GRIDBASE_H
#define LIBGEODECOMP_STORAGE_GRIDBASE_H

#include <libgeodecomp/geometry/coordbox.h>
#include <libgeodecomp/geometry/region.h>
#include <libgeodecomp/geometry/streak.h>

namespace LibGeoDecomp {

template<typename CELL, int DIMENSIONS>
class GridBase
{
public:
    typedef CELL CellType;
    const static int DIM = DIMENSIONS;

    virtual ~GridBase() {}

    virtual void resize(const CoordBox<DIM>&) = 0;
    virtual void set(const Coord<DIM>&, const CELL&) = 0;
    virtual void set(const Streak<DIM>&, const CELL*) = 0;
    virtual CELL get(const Coord<DIM>&) const = 0;
    virtual void get(const Streak<DIM>&, CELL*) const = 0;
    virtual void setEdge(const CELL&) = 0;
    virtual const CELL& getEdge() const = 0;
    virtual CoordBox<DIM> boundingBox() const = 0;

    virtual Region<DIM> boundingRegion() const
    {
        Region<DIM> r;
        r << boundingBox();
        return r;
    }

    Coord<DIM> dimensions() const { return boundingBox().dimensions; }

    bool operator==(const GridBase& other) const
    {
        if (boundingBox() != other.boundingBox()) {
            return false;
        }
        CoordBox<DIM> box = boundingBox();
        for (typename CoordBox<DIM>::Iterator i = box.begin(); i != box.end(); ++i) {
            if (!(get(*i) == other.get(*i))) {
                return false;
            }
        }
        return true;
    }
};

}

#endif
