// SerialBOVWriter<CELL>: writes one member of the grid as VisIt "Brick of
// Values" snapshots, a text header PREFIX.SSSSS.bov and a raw brick
// PREFIX.SSSSS.data (x fastest, native little-endian), at the initial step,
// every `period` steps and the last step (src/libgeodecomp/io/
// serialbovwriter.h:18-84, header and brick as in
// src/libgeodecomp/io/bovoutput.h:22-96: 1D and 2D grids are written as
// bricks with unit extent in the missing dimensions).  It only reads the
// grid through GridBase, so it serves any simulator, HipSimulator included.
#ifndef LIBGEODECOMP_IO_SERIALBOVWRITER_H
#define LIBGEODECOMP_IO_SERIALBOVWRITER_H

#include <libgeodecomp/geometry/coord.h>
#include <libgeodecomp/geometry/coordbox.h>
#include <libgeodecomp/geometry/streak.h>
#include <libgeodecomp/io/writer.h>
#include <libgeodecomp/storage/selector.h>

#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <vector>

namespace LibGeoDecomp {

template<typename CELL_TYPE>
class SerialBOVWriter : public Writer<CELL_TYPE>
{
public:
    typedef typename Writer<CELL_TYPE>::GridType GridType;
    static const int DIM = Writer<CELL_TYPE>::DIM;
    using Writer<CELL_TYPE>::period;
    using Writer<CELL_TYPE>::prefix;

    SerialBOVWriter(const Selector<CELL_TYPE>& selector, const std::string& prefix, const unsigned period,
                    const Coord<3>& brickletDim = Coord<3>()) :
        Writer<CELL_TYPE>(prefix, period),
        selector(selector),
        brickletDim(brickletDim)
    {}

    template<typename MEMBER>
    SerialBOVWriter(MEMBER CELL_TYPE:: *member, const std::string& prefix, const unsigned period,
                    const Coord<3>& brickletDim = Coord<3>()) :
        Writer<CELL_TYPE>(prefix, period),
        selector(member, prefix),
        brickletDim(brickletDim)
    {}

    Writer<CELL_TYPE> *clone() const { return new SerialBOVWriter(*this); }

    void stepFinished(const GridType& grid, unsigned step, WriterEvent event)
    {
        if (event == WRITER_STEP_FINISHED && step % period != 0) {
            return;
        }
        const std::string header = filename(step, "bov");
        const std::string data = filename(step, "data");
        writeHeader(header, data, step, grid.boundingBox());
        writeBrick(data, grid);
    }

    std::string filename(unsigned step, const std::string& suffix) const
    {
        std::ostringstream buf;
        buf << prefix << "." << std::setfill('0') << std::setw(5) << step << "." << suffix;
        return buf.str();
    }

private:
    Selector<CELL_TYPE> selector;
    Coord<3> brickletDim;

    void writeHeader(const std::string& header, const std::string& data, unsigned step,
                     const CoordBox<DIM>& box) const
    {
        Coord<3> size(1, 1, 1);
        Coord<3> origin(0, 0, 0);
        for (int d = 0; d < DIM; ++d) {
            size[d] = box.dimensions[d];
            origin[d] = box.origin[d];
        }
        const Coord<3> bricklets = (brickletDim == Coord<3>()) ? size : brickletDim;
        std::ofstream out(header.c_str());
        if (!out) {
            throw std::runtime_error("SerialBOVWriter: could not open " + header);
        }
        out << "TIME: " << step << "\n"
            << "DATA_FILE: " << data << "\n"
            << "DATA_SIZE: " << size.x() << " " << size.y() << " " << size.z() << "\n"
            << "DATA_FORMAT: " << selector.typeName() << "\n"
            << "VARIABLE: " << selector.name() << "\n"
            << "DATA_ENDIAN: LITTLE\n"
            << "BRICK_ORIGIN: " << origin.x() << " " << origin.y() << " " << origin.z() << "\n"
            << "BRICK_SIZE: " << size.x() << " " << size.y() << " " << size.z() << "\n"
            << "DIVIDE_BRICK: true\n"
            << "DATA_BRICKLETS: " << bricklets.x() << " " << bricklets.y() << " " << bricklets.z() << "\n"
            << "DATA_COMPONENTS: " << selector.arity() << "\n";
    }

    void writeBrick(const std::string& data, const GridType& grid) const
    {
        std::ofstream out(data.c_str(), std::ios::binary);
        if (!out.good()) {
            throw std::runtime_error("SerialBOVWriter: could not open " + data);
        }
        const CoordBox<DIM> box = grid.boundingBox();
        const int width = box.dimensions[0];
        std::vector<CELL_TYPE> row(width);
        std::vector<char> bytes(width * selector.sizeOfExternal());
        for (typename CoordBox<DIM>::StreakIterator i = box.beginStreak(); i != box.endStreak(); ++i) {
            const Streak<DIM> s = *i;
            grid.get(s, row.data());
            selector.copyMemberOut(row.data(), bytes.data(), (std::size_t)width);
            out.write(bytes.data(), (std::streamsize)bytes.size());
        }
    }
};

}

#endif
