// Selector<CELL>: names one member of a cell for output plugins, with the
// member's external type name and size (after
// src/libgeodecomp/storage/selector.h:189-300 and the type names of
// src/libgeodecomp/storage/filterbase.h:25-110: bool/char "BYTE", float
// "FLOAT", double "DOUBLE", int "INT", long "LONG"; other member types have
// no name and typeName() throws std::invalid_argument, as there).
#ifndef LIBGEODECOMP_STORAGE_SELECTOR_H
#define LIBGEODECOMP_STORAGE_SELECTOR_H

#include <cstddef>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>

namespace LibGeoDecomp {

namespace SelectorHelpers {

template<typename T>
struct TypeName
{
    static std::string get() { throw std::invalid_argument("no string representation known for member type"); }
};
template<> struct TypeName<bool> { static std::string get() { return "BYTE"; } };
template<> struct TypeName<char> { static std::string get() { return "BYTE"; } };
template<> struct TypeName<float> { static std::string get() { return "FLOAT"; } };
template<> struct TypeName<double> { static std::string get() { return "DOUBLE"; } };
template<> struct TypeName<int> { static std::string get() { return "INT"; } };
template<> struct TypeName<long> { static std::string get() { return "LONG"; } };

}

template<typename CELL>
class Selector
{
public:
    Selector() : size(0), typeNameOf(nullptr) {}

    template<typename MEMBER>
    Selector(MEMBER CELL:: *member, const std::string& variableName) :
        variableName(variableName),
        size(sizeof(MEMBER)),
        typeNameOf(&SelectorHelpers::TypeName<MEMBER>::get),
        extract([member](const CELL *cells, char *target, std::size_t count) {
            for (std::size_t i = 0; i < count; ++i) {
                std::memcpy(target + i * sizeof(MEMBER), &(cells[i].*member), sizeof(MEMBER));
            }
        })
    {}

    const std::string& name() const { return variableName; }
    std::string typeName() const
    {
        if (!typeNameOf) {
            throw std::logic_error("empty selector");
        }
        return typeNameOf();
    }
    int arity() const { return 1; }
    std::size_t sizeOfExternal() const { return size; }

    // the member of `count` consecutive cells, packed, into target
    void copyMemberOut(const CELL *cells, char *target, std::size_t count) const
    {
        if (!extract) {
            throw std::logic_error("empty selector");
        }
        extract(cells, target, count);
    }

private:
    std::string variableName;
    std::size_t size;
    std::string (*typeNameOf)();
    std::function<void(const CELL *, char *, std::size_t)> extract;
};

}

#endif
