// SimulationParameters: a named set of tunable parameters, each either an
// Interval of numbers or a DiscreteSet of values, that optimizers move
// through as plain doubles (getValue / setValue / +=) and simulation
// factories read back as typed values.  Interface and semantics after
// src/libgeodecomp/misc/simulationparameters.h:18-409:
//  * an Interval [min, max) with granularity g keeps an index in
//    [0, max - min - g], rounded down to multiples of g; its value is
//    min + index;
//  * a DiscreteSet keeps an index in [0, size - 1]; its value is the
//    element at that index;
//  * reading a parameter as a type it does not hold throws
//    std::logic_error.
#ifndef LIBGEODECOMP_MISC_SIMULATIONPARAMETERS_H
#define LIBGEODECOMP_MISC_SIMULATIONPARAMETERS_H

#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace LibGeoDecomp {

namespace SimulationParametersHelpers {

class OptimizableParameter
{
public:
    virtual ~OptimizableParameter() {}
    virtual double getMin() const = 0;
    virtual double getMax() const = 0;
    virtual double getGranularity() const = 0;
    virtual double getValue() const = 0;
    virtual void setValue(double newValue) = 0;
    virtual void operator+=(double step) = 0;
    virtual std::string toString() const = 0;
};

class Parameter : public OptimizableParameter
{
public:
    virtual Parameter *clone() const = 0;

    virtual operator std::string() const { throw std::logic_error("parameter is not a string"); }
    virtual operator bool() const { throw std::logic_error("parameter is not a bool"); }
    virtual operator int() const { throw std::logic_error("parameter is not an int"); }
    virtual operator double() const { throw std::logic_error("parameter is not a double"); }

    virtual void operator=(const std::string&) { throw std::logic_error("parameter is not a string"); }
    virtual void operator=(const bool&) { throw std::logic_error("parameter is not a bool"); }
    virtual void operator=(const int&) { throw std::logic_error("parameter is not an int"); }
    virtual void operator=(const double&) { throw std::logic_error("parameter is not a double"); }

    virtual bool operator==(const std::string&) const { return false; }
    virtual bool operator==(const bool&) const { return false; }
    virtual bool operator==(const int&) const { return false; }
    virtual bool operator==(const double&) const { return false; }
    bool operator==(const char *s) const { return *this == std::string(s); }
    template<typename T>
    bool operator!=(const T& o) const { return !(*this == o); }

protected:
    // clamp an index into [getMin(), getMax() - getGranularity()]
    double clampIndex(double index) const
    {
        if (index < getMin()) {
            index = getMin();
        }
        if (index >= getMax()) {
            index = getMax() - getGranularity();
        }
        return index;
    }
};

template<typename VALUE>
class TypedParameter : public Parameter
{
public:
    using Parameter::operator=;
    using Parameter::operator==;

    explicit TypedParameter(const VALUE& v) : current(v) {}

    virtual operator VALUE() const { return current; }
    virtual void operator=(const VALUE& v) { current = v; }
    virtual bool operator==(const VALUE& v) const { return current == v; }

protected:
    VALUE current;
};

template<typename VALUE>
class Interval : public TypedParameter<VALUE>
{
public:
    Interval(const VALUE minimum, const VALUE maximum, const double granularity = 1) :
        TypedParameter<VALUE>(minimum), minimum(minimum), maximum(maximum), granularity(granularity), index(0)
    {}

    Parameter *clone() const { return new Interval(*this); }
    double getMin() const { return 0; }
    double getMax() const { return (double)(maximum - minimum); }
    double getGranularity() const { return granularity; }
    double getValue() const { return (double)index; }

    void setValue(double v) { place(v); }
    void operator+=(double step) { place((double)index + step); }

    std::string toString() const
    {
        std::ostringstream s;
        s << "Interval([" << minimum << ", " << maximum << "], " << index << ")";
        return s.str();
    }

private:
    VALUE minimum;
    VALUE maximum;
    double granularity;
    VALUE index;

    void place(double v)
    {
        if (granularity > 0.0) {
            v = (double)(long)(v / granularity) * granularity;
        }
        index = (VALUE)this->clampIndex(v);
        this->current = minimum + index;
    }
};

template<typename VALUE>
class DiscreteSet : public TypedParameter<VALUE>
{
public:
    explicit DiscreteSet(const std::vector<VALUE>& elements) :
        TypedParameter<VALUE>(elements.at(0)), elements(elements), index(0)
    {}

    Parameter *clone() const { return new DiscreteSet(*this); }
    double getMin() const { return 0; }
    double getMax() const { return (double)elements.size(); }
    double getGranularity() const { return 1; }
    double getValue() const { return (double)index; }

    void setValue(double v) { place(v); }
    void operator+=(double step) { place((double)index + step); }

    std::string toString() const
    {
        std::ostringstream s;
        s << "DiscreteSet([";
        for (std::size_t i = 0; i < elements.size(); ++i) {
            s << (i ? ", " : "") << elements[i];
        }
        s << "], " << index << ")";
        return s.str();
    }

private:
    std::vector<VALUE> elements;
    int index;

    void place(double v)
    {
        index = (int)this->clampIndex(v);
        this->current = elements[index];
    }
};

}

class SimulationParameters
{
public:
    typedef SimulationParametersHelpers::Parameter Parameter;

    SimulationParameters() {}
    SimulationParameters(const SimulationParameters& o) : names(o.names)
    {
        for (std::size_t i = 0; i < o.parameters.size(); ++i) {
            parameters.emplace_back(o.parameters[i]->clone());
        }
    }
    SimulationParameters& operator=(const SimulationParameters& o)
    {
        if (this != &o) {
            SimulationParameters copy(o);
            names.swap(copy.names);
            parameters.swap(copy.parameters);
        }
        return *this;
    }

    template<typename VALUE>
    void addParameter(const std::string& name, const VALUE& minimum, const VALUE& maximum,
                      const double granularity = 1.0)
    {
        add(name, new SimulationParametersHelpers::Interval<VALUE>(minimum, maximum, granularity));
    }

    template<typename VALUE>
    void addParameter(const std::string& name, const std::vector<VALUE>& elements)
    {
        add(name, new SimulationParametersHelpers::DiscreteSet<VALUE>(elements));
    }

    template<typename VALUE>
    void replaceParameter(const std::string& name, const VALUE& minimum, const VALUE& maximum,
                          const double granularity = 1.0)
    {
        parameters[indexOf(name)].reset(new SimulationParametersHelpers::Interval<VALUE>(minimum, maximum,
                                                                                         granularity));
    }

    template<typename VALUE>
    void replaceParameter(const std::string& name, const std::vector<VALUE>& elements)
    {
        parameters[indexOf(name)].reset(new SimulationParametersHelpers::DiscreteSet<VALUE>(elements));
    }

    Parameter& operator[](const std::string& name) { return *parameters[indexOf(name)]; }
    const Parameter& operator[](const std::string& name) const { return *parameters[indexOf(name)]; }
    Parameter& operator[](std::size_t i) { return *parameters.at(i); }
    const Parameter& operator[](std::size_t i) const { return *parameters.at(i); }

    std::size_t size() const { return parameters.size(); }
    // the name of parameter i (parameters are indexed in the order added)
    std::string name(std::size_t i) const
    {
        for (std::map<std::string, std::size_t>::const_iterator n = names.begin(); n != names.end(); ++n) {
            if (n->second == i) {
                return n->first;
            }
        }
        throw std::out_of_range("no parameter with this index");
    }
    bool contains(const std::string& name) const { return names.count(name) != 0; }

    std::string toString() const
    {
        std::ostringstream s;
        s << "SimulationParameters(";
        for (std::map<std::string, std::size_t>::const_iterator i = names.begin(); i != names.end(); ++i) {
            s << (i == names.begin() ? "" : ", ") << i->first << ": " << parameters[i->second]->toString();
        }
        s << ")";
        return s.str();
    }

private:
    std::map<std::string, std::size_t> names;
    std::vector<std::unique_ptr<Parameter> > parameters;

    void add(const std::string& name, Parameter *p)
    {
        if (names.count(name)) {
            delete p;
            throw std::invalid_argument("parameter " + name + " exists already");
        }
        names[name] = parameters.size();
        parameters.emplace_back(p);
    }

    std::size_t indexOf(const std::string& name) const
    {
        std::map<std::string, std::size_t>::const_iterator i = names.find(name);
        if (i == names.end()) {
            throw std::invalid_argument("no parameter named " + name);
        }
        return i->second;
    }
};

}

#endif
