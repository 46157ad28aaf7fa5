// Coord<DIM>: integer lattice coordinate, x fastest.
// Interface after src/libgeodecomp/geometry/coord.h (x()/y()/z(), operator[],
// prod(), toIndex() with x fastest, coord.h:427-430 and :710-716).
#ifndef LIBGEODECOMP_GEOMETRY_COORD_H
#define LIBGEODECOMP_GEOMETRY_COORD_H

#include <libgeodecomp/geometry/fixedcoord.h>
#include <libgeodecomp/misc/hostdevice.h>

#include <cstddef>
#include <ostream>
#include <string>
#include <sstream>

namespace LibGeoDecomp {

template<int DIM>
class Coord
{
public:
    static const int DIMENSIONS = DIM;

    __host__ __device__ explicit Coord(int x = 0, int y = 0, int z = 0)
    {
        int v[3] = {x, y, z};
        for (int d = 0; d < DIM; ++d) {
            c[d] = v[d];
        }
    }

    // the offset of a FixedCoord, truncated to DIM components
    template<int X, int Y, int Z>
    __host__ __device__ explicit Coord(FixedCoord<X, Y, Z>)
    {
        const int v[3] = {X, Y, Z};
        for (int d = 0; d < DIM; ++d) {
            c[d] = v[d];
        }
    }

    __host__ __device__ static Coord diagonal(int n)
    {
        Coord r;
        for (int d = 0; d < DIM; ++d) {
            r.c[d] = n;
        }
        return r;
    }

    __host__ __device__ int& x() { return c[0]; }
    __host__ __device__ int x() const { return c[0]; }
    __host__ __device__ int& y() { static_assert(DIM >= 2, "no y in 1D"); return c[1]; }
    __host__ __device__ int y() const { static_assert(DIM >= 2, "no y in 1D"); return c[1]; }
    __host__ __device__ int& z() { static_assert(DIM >= 3, "no z below 3D"); return c[2]; }
    __host__ __device__ int z() const { static_assert(DIM >= 3, "no z below 3D"); return c[2]; }

    __host__ __device__ int& operator[](int d) { return c[d]; }
    __host__ __device__ int operator[](int d) const { return c[d]; }

    __host__ __device__ Coord operator+(const Coord& o) const { Coord r; for (int d = 0; d < DIM; ++d) r.c[d] = c[d] + o.c[d]; return r; }
    __host__ __device__ Coord operator-(const Coord& o) const { Coord r; for (int d = 0; d < DIM; ++d) r.c[d] = c[d] - o.c[d]; return r; }
    __host__ __device__ Coord operator-() const { Coord r; for (int d = 0; d < DIM; ++d) r.c[d] = -c[d]; return r; }
    __host__ __device__ Coord operator*(int s) const { Coord r; for (int d = 0; d < DIM; ++d) r.c[d] = c[d] * s; return r; }
    __host__ __device__ Coord& operator+=(const Coord& o) { for (int d = 0; d < DIM; ++d) c[d] += o.c[d]; return *this; }
    __host__ __device__ Coord& operator-=(const Coord& o) { for (int d = 0; d < DIM; ++d) c[d] -= o.c[d]; return *this; }

    __host__ __device__ bool operator==(const Coord& o) const { for (int d = 0; d < DIM; ++d) if (c[d] != o.c[d]) return false; return true; }
    __host__ __device__ bool operator!=(const Coord& o) const { return !(*this == o); }
    // ordering: slowest dimension first (z, then y, then x)
    __host__ __device__ bool operator<(const Coord& o) const
    {
        for (int d = DIM - 1; d >= 0; --d) {
            if (c[d] != o.c[d]) {
                return c[d] < o.c[d];
            }
        }
        return false;
    }

    __host__ __device__ long prod() const { long p = 1; for (int d = 0; d < DIM; ++d) p *= c[d]; return p; }
    __host__ __device__ int sum() const { int s = 0; for (int d = 0; d < DIM; ++d) s += c[d]; return s; }

    // linear index of this coordinate inside a box of the given extent, x fastest
    __host__ __device__ std::size_t toIndex(const Coord& dim) const
    {
        std::size_t i = 0;
        for (int d = DIM - 1; d >= 0; --d) {
            i = i * static_cast<std::size_t>(dim.c[d]) + static_cast<std::size_t>(c[d]);
        }
        return i;
    }

    std::string toString() const
    {
        std::ostringstream s;
        s << "(";
        for (int d = 0; d < DIM; ++d) {
            s << (d ? ", " : "") << c[d];
        }
        s << ")";
        return s.str();
    }

private:
    int c[DIM > 0 ? DIM : 1];
};

template<int DIM>
std::ostream& operator<<(std::ostream& os, const Coord<DIM>& c)
{
    return os << c.toString();
}

}

#endif
