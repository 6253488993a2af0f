// nnp::ElementMap (ElementMap.h of the reference): element symbols of a
// model, sorted by atomic number, and their indices.
#ifndef NNP_ELEMENTMAP_H
#define NNP_ELEMENTMAP_H

#include <cstddef>
#include <map>
#include <string>
#include <vector>

namespace nnp
{

class ElementMap
{
public:
    std::string operator[](std::size_t const index) const { return symbol(index); }
    std::size_t operator[](std::string const symbol) const { return index(symbol); }
    std::size_t size() const { return forward.size(); }
    std::string getElementsString() const;
    std::size_t index(std::string const& symbol) const;
    std::string symbol(std::size_t const index) const;
    std::size_t atomicNumber(std::size_t index) const
    {
        return atomicNumber(symbol(index));
    }
    // Registers the symbols of an "elements" line, sorted by atomic number
    // (ElementMap.cpp:36-54); returns the element count.
    std::size_t registerElements(std::string const& elementLine);
    void deregisterElements();
    static std::string symbolFromAtomicNumber(std::size_t const atomicNumber);
    static std::size_t atomicNumber(std::string const& symbol);
    std::vector<std::string> info() const;

private:
    std::vector<std::string> forward;
    std::map<std::string, std::size_t> reverse;
};

}  // namespace nnp

#endif
