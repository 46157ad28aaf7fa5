// MPIIO<CELL>: reads and writes the reference's MPI-IO grid snapshot format
// (src/libgeodecomp/io/mpiio.h:26-233) with plain file I/O -- one process
// drives one GPU here, so there is no communicator and the whole file is
// written and read by that process.  The bytes are the reference's:
//   header  Coord<DIM> dimensions (DIM x int32), unsigned step,
//           unsigned maxSteps, the edge cell (sizeof(CELL) bytes)
//   body    the cells of the bounding box, x fastest (Coord::toIndex order),
//           sizeof(CELL) bytes each
// (the reference's lengths are the extents of the MPI datatypes of Coord,
// unsigned and the cell, which are these sizes for the automatically
// generated typemaps of plain structs).  Regions are read and written
// streak by streak at offset header + index * sizeof(CELL), torus
// coordinates normalised first, as the reference does (:45-56, :112-127).
#ifndef LIBGEODECOMP_IO_MPIIO_H
#define LIBGEODECOMP_IO_MPIIO_H

#include <libgeodecomp/geometry/coord.h>
#include <libgeodecomp/geometry/region.h>
#include <libgeodecomp/misc/apitraits.h>
#include <libgeodecomp/storage/gridbase.h>

#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace LibGeoDecomp {

template<typename CELL_TYPE, typename TOPOLOGY = typename APITraits::SelectTopology<CELL_TYPE>::Value>
class MPIIO
{
public:
    static_assert(std::is_trivially_copyable<CELL_TYPE>::value,
                  "snapshots store cells as their bytes: the cell must be trivially copyable");

    template<int DIM>
    static std::size_t headerLength()
    {
        return DIM * sizeof(int32_t) + 2 * sizeof(uint32_t) + sizeof(CELL_TYPE);
    }

    template<typename GRID_TYPE, int DIM>
    void readRegion(GRID_TYPE *grid, const std::string& filename, const Region<DIM>& region)
    {
        std::ifstream file(filename.c_str(), std::ios::binary);
        if (!file) {
            throw std::runtime_error("MPIIO: could not open " + filename + " for reading");
        }
        const Coord<DIM> dimensions = readDimensions<DIM>(file, filename);
        // the edge cell is the last element of the header
        file.seekg((std::streamoff)(headerLength<DIM>() - sizeof(CELL_TYPE)));
        CELL_TYPE edge;
        readBytes(file, &edge, sizeof(CELL_TYPE), filename);
        grid->setEdge(edge);
        std::vector<CELL_TYPE> line;
        const std::vector<Streak<DIM> > streaks = region.streaks();
        for (std::size_t s = 0; s < streaks.size(); ++s) {
            const Coord<DIM> coord = TOPOLOGY::normalize(streaks[s].origin, dimensions);
            const int length = streaks[s].endX - streaks[s].origin.x();
            line.resize(length);
            file.seekg((std::streamoff)offset(coord, dimensions));
            readBytes(file, line.data(), sizeof(CELL_TYPE) * length, filename);
            grid->set(streaks[s], line.data());
        }
    }

    template<int DIM>
    void readMetadata(Coord<DIM> *dimensions, unsigned *step, unsigned *maxSteps, const std::string& filename)
    {
        std::ifstream file(filename.c_str(), std::ios::binary);
        if (!file) {
            throw std::runtime_error("MPIIO: could not open " + filename + " for reading");
        }
        *dimensions = readDimensions<DIM>(file, filename);
        uint32_t s, m;
        readBytes(file, &s, sizeof(s), filename);
        readBytes(file, &m, sizeof(m), filename);
        *step = s;
        *maxSteps = m;
    }

    // writes the header and the region's cells; a file that already exists
    // is overwritten in place (its other bytes are kept, as MPI_File_open
    // without MPI_MODE_EXCL keeps them)
    template<typename GRID_TYPE, int DIM>
    void writeRegion(const GRID_TYPE& grid, const Coord<DIM>& dimensions, unsigned step, unsigned maxSteps,
                     const std::string& filename, const Region<DIM>& region)
    {
        std::fstream file(filename.c_str(), std::ios::binary | std::ios::in | std::ios::out);
        if (!file) {
            file.clear();
            file.open(filename.c_str(), std::ios::binary | std::ios::out);
        }
        if (!file) {
            throw std::runtime_error("MPIIO: could not open " + filename + " for writing");
        }
        int32_t dims[DIM];
        for (int d = 0; d < DIM; ++d) {
            dims[d] = dimensions[d];
        }
        const uint32_t s = step, m = maxSteps;
        const CELL_TYPE edge = grid.getEdge();
        file.seekp(0);
        file.write(reinterpret_cast<const char *>(dims), sizeof(dims));
        file.write(reinterpret_cast<const char *>(&s), sizeof(s));
        file.write(reinterpret_cast<const char *>(&m), sizeof(m));
        file.write(reinterpret_cast<const char *>(&edge), sizeof(CELL_TYPE));
        std::vector<CELL_TYPE> line;
        const std::vector<Streak<DIM> > streaks = region.streaks();
        for (std::size_t k = 0; k < streaks.size(); ++k) {
            const Coord<DIM> coord = TOPOLOGY::normalize(streaks[k].origin, dimensions);
            const int length = streaks[k].endX - streaks[k].origin.x();
            line.resize(length);
            grid.get(streaks[k], line.data());
            file.seekp((std::streamoff)offset(coord, dimensions));
            file.write(reinterpret_cast<const char *>(line.data()), (std::streamsize)(sizeof(CELL_TYPE) * length));
        }
        if (!file) {
            throw std::runtime_error("MPIIO: writing " + filename + " failed");
        }
    }

private:
    template<int DIM>
    static std::size_t offset(const Coord<DIM>& c, const Coord<DIM>& dimensions)
    {
        return headerLength<DIM>() + c.toIndex(dimensions) * sizeof(CELL_TYPE);
    }

    static void readBytes(std::ifstream& file, void *dst, std::size_t n, const std::string& filename)
    {
        file.read(reinterpret_cast<char *>(dst), (std::streamsize)n);
        if (!file) {
            throw std::runtime_error("MPIIO: " + filename + " is shorter than its header says");
        }
    }

    template<int DIM>
    static Coord<DIM> readDimensions(std::ifstream& file, const std::string& filename)
    {
        int32_t dims[DIM];
        readBytes(file, dims, sizeof(dims), filename);
        Coord<DIM> ret;
        for (int d = 0; d < DIM; ++d) {
            ret[d] = dims[d];
        }
        return ret;
    }
};

}

#endif
