// Struct-of-arrays registration and accessors for user cells, usable on the
// host and inside HIP kernels.
//
// User interface kept from libflatarray (the reference's vendored
// lib/libflatarray/include/libflatarray/macros.hpp:81-127, :130-388):
//   LIBFLATARRAY_REGISTER_SOA(Cell, ((double)(temp))((int)(flags))((float)(f)(19)))
//   LIBFLATARRAY_ACCESS                       (inside a cell with private members)
//   accessor.temp()                           member reference (arrays: f<i>(), f()[i])
//   accessor.index(), ++accessor.index(), accessor += n
//   accessor[LibFlatArray::coord<X, Y, Z>()]  accessor at a fixed offset
//   accessor << cell, accessor >> cell, (Cell) accessor
//   LibFlatArray::api_traits::has_default_*_sizes*   (marker classes)
//
// Storage semantics kept (detail/macros.hpp:99-106): member-major planes,
// member m of element i at
//     data + DIM_PROD * offset(m) + i * sizeof(m)
// with offset(m) the summed sizes of the members registered before m (an
// array member of n elements is n consecutive planes) and i = z * DIM_X *
// DIM_Y + y * DIM_X + x.
//
// What differs, by design for the GPU: the grid extents DIM_X, DIM_Y and
// DIM_PROD are run-time values carried by the accessor, not template
// parameters chosen from a list of size buckets (libflatarray instantiates
// every bucket so the CPU compiler can fold the address arithmetic; a HIP
// kernel keeps the extents in scalar registers at no cost, and the grid is
// sized to the problem and aligned for coalesced rows instead of rounded
// up to a bucket).  The bucket marker classes are accepted and ignored.
#ifndef LIBFLATARRAY_FLAT_ARRAY_HPP
#define LIBFLATARRAY_FLAT_ARRAY_HPP

#include <libgeodecomp/misc/hostdevice.h>

#include <cstddef>

namespace LibFlatArray {

// compile-time offset of an accessor (libflatarray/coord.hpp)
template<long X, long Y, long Z>
class coord
{
public:
    static const long DIM_X = X;
    static const long DIM_Y = Y;
    static const long DIM_Z = Z;

    __host__ __device__ inline coord() {}
};

// Marker classes a cell's API may derive from to pick libflatarray's size
// buckets (api_traits.hpp:67-140).  Grids here are sized at run time, so
// they carry no meaning on this path.
class api_traits
{
public:
    class has_default_1d_sizes {};
    class has_default_2d_sizes {};
    class has_default_2d_sizes_uniform {};
    class has_default_3d_sizes {};
    class has_default_3d_sizes_uniform {};
    template<typename SIZES>
    class has_custom_sizes {};
};

// Per-cell member table, specialised by LIBFLATARRAY_REGISTER_SOA.
template<typename CELL>
class soa_traits;

// Position of one element in a member-major grid: the grid's first byte,
// the element index and the extents that turn offsets into indices.
class soa_position
{
public:
    __host__ __device__ inline soa_position(char *data, long index, long dimX, long dimXY, long dimProd) :
        my_data(data), my_index(index), my_dim_x(dimX), my_dim_xy(dimXY), my_dim_prod(dimProd)
    {}

    // address of element my_index of the plane that starts at byte
    // dimProd * offset, for a member of `size` bytes
    __host__ __device__ inline char *member_address(long offset, long size) const
    {
        return my_data + my_dim_prod * offset + my_index * size;
    }

    char *my_data;
    long my_index;
    long my_dim_x;
    long my_dim_xy;
    long my_dim_prod;
};

// Array member proxy: name()[i] on an accessor.
template<typename T>
class soa_member_array
{
public:
    __host__ __device__ inline soa_member_array(char *first, long planeBytes) : first(first), planeBytes(planeBytes) {}

    __host__ __device__ inline T& operator[](long i) const { return *(T *)(first + i * planeBytes); }

private:
    char *first;
    long planeBytes;
};

// Accessor of one element of a registered cell type: the member functions
// named after the cell's members come from soa_traits<CELL>::members.
template<typename CELL>
class soa_accessor : public soa_traits<CELL>::template members<soa_position>
{
public:
    typedef CELL element_type;
    typedef typename soa_traits<CELL>::template members<soa_position> base_type;

    __host__ __device__ inline soa_accessor(char *data, long index, long dimX, long dimXY, long dimProd) :
        base_type(data, index, dimX, dimXY, dimProd)
    {}

    __host__ __device__ inline long gen_index(long x, long y, long z) const
    {
        return z * this->my_dim_xy + y * this->my_dim_x + x;
    }

    __host__ __device__ inline long& index() { return this->my_index; }
    __host__ __device__ inline const long& index() const { return this->my_index; }
    __host__ __device__ inline char *data() const { return this->my_data; }
    __host__ __device__ inline long dim_x() const { return this->my_dim_x; }
    __host__ __device__ inline long dim_xy() const { return this->my_dim_xy; }
    __host__ __device__ inline long dim_prod() const { return this->my_dim_prod; }

    __host__ __device__ inline void operator+=(long offset) { this->my_index += offset; }
    __host__ __device__ inline void operator++() { ++this->my_index; }

    template<long X, long Y, long Z>
    __host__ __device__ inline soa_accessor operator[](coord<X, Y, Z>) const
    {
        return soa_accessor(this->my_data, this->my_index + gen_index(X, Y, Z), this->my_dim_x, this->my_dim_xy,
                            this->my_dim_prod);
    }

    __host__ __device__ inline operator CELL() const
    {
        CELL cell;
        soa_traits<CELL>::gather(&cell, *this);
        return cell;
    }

    __host__ __device__ inline void operator<<(const CELL& cell) { soa_traits<CELL>::scatter(cell, this); }
    __host__ __device__ inline void operator>>(CELL& cell) const { soa_traits<CELL>::gather(&cell, *this); }
    __host__ __device__ inline soa_accessor& operator=(const CELL& cell)
    {
        soa_traits<CELL>::scatter(cell, this);
        return *this;
    }
};

}

// ---------------------------------------------------------------------
// Registration macro.  A member is ((type)(name)) or ((type)(name)(count));
// the preprocessor helpers below take one such element apart by counting
// its parenthesised groups (no commas, so types may carry :: and <N>).
// ---------------------------------------------------------------------
#define LGD_SOA_CAT(a, b) LGD_SOA_CAT_I(a, b)
#define LGD_SOA_CAT_I(a, b) a##b

// type of (t)(n) / (t)(n)(c)
#define LGD_SOA_TYPE(elem) LGD_SOA_CAT(LGD_SOA_TYPE_A elem, _END)
#define LGD_SOA_TYPE_A(x) x LGD_SOA_TYPE_B
#define LGD_SOA_TYPE_B(x) LGD_SOA_TYPE_C
#define LGD_SOA_TYPE_C(x) LGD_SOA_TYPE_D
#define LGD_SOA_TYPE_C_END
#define LGD_SOA_TYPE_D_END
// name
#define LGD_SOA_NAME(elem) LGD_SOA_CAT(LGD_SOA_NAME_A elem, _END)
#define LGD_SOA_NAME_A(x) LGD_SOA_NAME_B
#define LGD_SOA_NAME_B(x) x LGD_SOA_NAME_C
#define LGD_SOA_NAME_C(x) LGD_SOA_NAME_D
#define LGD_SOA_NAME_C_END
#define LGD_SOA_NAME_D_END
// element count (1 for a scalar member)
#define LGD_SOA_COUNT(elem) LGD_SOA_CAT(LGD_SOA_COUNT_A elem, _END)
#define LGD_SOA_COUNT_A(x) LGD_SOA_COUNT_B
#define LGD_SOA_COUNT_B(x) LGD_SOA_COUNT_C
#define LGD_SOA_COUNT_C(x) x LGD_SOA_COUNT_D
#define LGD_SOA_COUNT_C_END 1
#define LGD_SOA_COUNT_D_END
// 0 for a scalar member, 1 for an array member
#define LGD_SOA_ISARRAY(elem) LGD_SOA_CAT(LGD_SOA_ISARRAY_A elem, _END)
#define LGD_SOA_ISARRAY_A(x) LGD_SOA_ISARRAY_B
#define LGD_SOA_ISARRAY_B(x) LGD_SOA_ISARRAY_C
#define LGD_SOA_ISARRAY_C(x) LGD_SOA_ISARRAY_D
#define LGD_SOA_ISARRAY_C_END 0
#define LGD_SOA_ISARRAY_D_END 1

// One pass over the member sequence per generated piece (a pass expands
// a fixed per-member macro; the sequence ((..))((..)).. is walked by two
// alternating macros and the trailing one is pasted into an empty _END).
#define LGD_SOA_PASS_LAYOUT(seq) LGD_SOA_CAT(LGD_SOA_LAYOUT_A seq, _END)
#define LGD_SOA_LAYOUT_A(elem) LGD_SOA_LAYOUT_ONE(elem) LGD_SOA_LAYOUT_B
#define LGD_SOA_LAYOUT_B(elem) LGD_SOA_LAYOUT_ONE(elem) LGD_SOA_LAYOUT_A
#define LGD_SOA_LAYOUT_A_END
#define LGD_SOA_LAYOUT_B_END
#define LGD_SOA_LAYOUT_ONE(elem) char LGD_SOA_NAME(elem)[sizeof(LGD_SOA_TYPE(elem)) * (LGD_SOA_COUNT(elem))];

#define LGD_SOA_PASS_NUMBER(seq) LGD_SOA_CAT(LGD_SOA_NUMBER_A seq, _END)
#define LGD_SOA_NUMBER_A(elem) +1 LGD_SOA_NUMBER_B
#define LGD_SOA_NUMBER_B(elem) +1 LGD_SOA_NUMBER_A
#define LGD_SOA_NUMBER_A_END
#define LGD_SOA_NUMBER_B_END

#define LGD_SOA_PASS_ACCESS(seq) LGD_SOA_CAT(LGD_SOA_ACCESS_A seq, _END)
#define LGD_SOA_ACCESS_A(elem) LGD_SOA_ACCESS_ONE(elem) LGD_SOA_ACCESS_B
#define LGD_SOA_ACCESS_B(elem) LGD_SOA_ACCESS_ONE(elem) LGD_SOA_ACCESS_A
#define LGD_SOA_ACCESS_A_END
#define LGD_SOA_ACCESS_B_END
#define LGD_SOA_ACCESS_ONE(elem) LGD_SOA_CAT(LGD_SOA_ACCESS_KIND_, LGD_SOA_ISARRAY(elem))(LGD_SOA_TYPE(elem), \
                                                                                        LGD_SOA_NAME(elem))
#define LGD_SOA_ACCESS_KIND_0(T, NAME)                                                                          \
    __host__ __device__ inline T& NAME()                                                                        \
    {                                                                                                           \
        return *(T *)this->member_address((long)offsetof(lgd_soa_layout, NAME), (long)sizeof(T));               \
    }                                                                                                           \
    __host__ __device__ inline const T& NAME() const                                                            \
    {                                                                                                           \
        return *(const T *)this->member_address((long)offsetof(lgd_soa_layout, NAME), (long)sizeof(T));         \
    }
#define LGD_SOA_ACCESS_KIND_1(T, NAME)                                                                          \
    template<int LGD_SOA_I>                                                                                     \
    __host__ __device__ inline T& NAME() const                                                                  \
    {                                                                                                           \
        return *(T *)this->member_address((long)offsetof(lgd_soa_layout, NAME) + LGD_SOA_I * (long)sizeof(T),  \
                                          (long)sizeof(T));                                                     \
    }                                                                                                           \
    __host__ __device__ inline ::LibFlatArray::soa_member_array<T> NAME() const                                 \
    {                                                                                                           \
        return ::LibFlatArray::soa_member_array<T>(                                                            \
            this->member_address((long)offsetof(lgd_soa_layout, NAME), (long)sizeof(T)),                      \
            this->my_dim_prod * (long)sizeof(T));                                                               \
    }

#define LGD_SOA_PASS_GATHER(seq) LGD_SOA_CAT(LGD_SOA_GATHER_A seq, _END)
#define LGD_SOA_GATHER_A(elem) LGD_SOA_GATHER_ONE(elem) LGD_SOA_GATHER_B
#define LGD_SOA_GATHER_B(elem) LGD_SOA_GATHER_ONE(elem) LGD_SOA_GATHER_A
#define LGD_SOA_GATHER_A_END
#define LGD_SOA_GATHER_B_END
#define LGD_SOA_GATHER_ONE(elem) LGD_SOA_CAT(LGD_SOA_GATHER_KIND_, LGD_SOA_ISARRAY(elem))(LGD_SOA_NAME(elem), \
                                                                                          LGD_SOA_COUNT(elem))
#define LGD_SOA_GATHER_KIND_0(NAME, N) cell->NAME = acc.NAME();
#define LGD_SOA_GATHER_KIND_1(NAME, N)                                                                          \
    for (int lgd_i = 0; lgd_i < (N); ++lgd_i) {                                                                 \
        cell->NAME[lgd_i] = acc.NAME()[lgd_i];                                                                  \
    }

#define LGD_SOA_PASS_SCATTER(seq) LGD_SOA_CAT(LGD_SOA_SCATTER_A seq, _END)
#define LGD_SOA_SCATTER_A(elem) LGD_SOA_SCATTER_ONE(elem) LGD_SOA_SCATTER_B
#define LGD_SOA_SCATTER_B(elem) LGD_SOA_SCATTER_ONE(elem) LGD_SOA_SCATTER_A
#define LGD_SOA_SCATTER_A_END
#define LGD_SOA_SCATTER_B_END
#define LGD_SOA_SCATTER_ONE(elem) LGD_SOA_CAT(LGD_SOA_SCATTER_KIND_, LGD_SOA_ISARRAY(elem))(LGD_SOA_NAME(elem), \
                                                                                            LGD_SOA_COUNT(elem))
#define LGD_SOA_SCATTER_KIND_0(NAME, N) acc->NAME() = cell.NAME;
#define LGD_SOA_SCATTER_KIND_1(NAME, N)                                                                         \
    for (int lgd_i = 0; lgd_i < (N); ++lgd_i) {                                                                 \
        static_cast<const ACC *>(acc)->NAME()[lgd_i] = cell.NAME[lgd_i];                                        \
    }

// Host-side table of the members: byte offset of the member's first plane
// per element of DIM_PROD, its bytes per element, and its offset in the cell.
#define LGD_SOA_PASS_TABLE(seq) LGD_SOA_CAT(LGD_SOA_TABLE_A seq, _END)
#define LGD_SOA_TABLE_A(elem) LGD_SOA_TABLE_ONE(elem) LGD_SOA_TABLE_B
#define LGD_SOA_TABLE_B(elem) LGD_SOA_TABLE_ONE(elem) LGD_SOA_TABLE_A
#define LGD_SOA_TABLE_A_END
#define LGD_SOA_TABLE_B_END
#define LGD_SOA_TABLE_ONE(elem)                                                                                 \
    f((std::size_t)offsetof(lgd_soa_layout, LGD_SOA_NAME(elem)), sizeof(LGD_SOA_TYPE(elem)),                   \
      (std::size_t)(LGD_SOA_COUNT(elem)));

#define LIBFLATARRAY_REGISTER_SOA(CELL_TYPE, CELL_MEMBERS)                                                      \
    namespace LibFlatArray {                                                                                    \
    template<>                                                                                                  \
    class soa_traits<CELL_TYPE>                                                                                 \
    {                                                                                                           \
    public:                                                                                                     \
        typedef CELL_TYPE cell_type;                                                                            \
        /* one char array per member in registration order: offsetof gives the                            \
           summed sizes of the members before it, sizeof the bytes per element */                         \
        struct lgd_soa_layout                                                                                   \
        {                                                                                                       \
            LGD_SOA_PASS_LAYOUT(CELL_MEMBERS)                                                                   \
        };                                                                                                      \
        static const int MEMBERS = 0 LGD_SOA_PASS_NUMBER(CELL_MEMBERS);                                         \
        static const std::size_t BYTES_PER_CELL = sizeof(lgd_soa_layout);                                      \
                                                                                                                \
        template<typename BASE>                                                                                 \
        class members : public BASE                                                                             \
        {                                                                                                       \
        public:                                                                                                 \
            __host__ __device__ inline members(char *data, long index, long dimX, long dimXY, long dimProd) :   \
                BASE(data, index, dimX, dimXY, dimProd)                                                         \
            {}                                                                                                  \
            LGD_SOA_PASS_ACCESS(CELL_MEMBERS)                                                                   \
        };                                                                                                      \
                                                                                                                \
        template<typename ACC>                                                                                  \
        __host__ __device__ static inline void gather(CELL_TYPE *cell, const ACC& acc)                          \
        {                                                                                                       \
            LGD_SOA_PASS_GATHER(CELL_MEMBERS)                                                                   \
        }                                                                                                       \
                                                                                                                \
        template<typename ACC>                                                                                  \
        __host__ __device__ static inline void scatter(const CELL_TYPE& cell, ACC *acc)                         \
        {                                                                                                       \
            LGD_SOA_PASS_SCATTER(CELL_MEMBERS)                                                                  \
        }                                                                                                       \
                                                                                                                \
        /* f(plane offset per element, bytes per element, elements) per member */                         \
        template<typename F>                                                                                    \
        static void each_member(F f)                                                                            \
        {                                                                                                       \
            LGD_SOA_PASS_TABLE(CELL_MEMBERS)                                                                    \
        }                                                                                                       \
    };                                                                                                          \
    }

// Gives the registration access to private members of a cell.
#define LIBFLATARRAY_ACCESS                                                                                     \
    template<typename LGD_SOA_CELL>                                                                             \
    friend class ::LibFlatArray::soa_traits;

#endif
