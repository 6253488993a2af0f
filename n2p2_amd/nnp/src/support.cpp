// utility.h, Log.h, ElementMap.h and Atom.h of the nnp layer.
#include "Atom.h"
#include "ElementMap.h"
#include "Log.h"
#include "utility.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace nnp
{

// ------------------------------------------------------------- utility --

std::vector<std::string> split(std::string const& input, char delimiter)
{
    std::vector<std::string> out;
    std::string item;
    std::istringstream in(input);
    while (std::getline(in, item, delimiter)) out.push_back(item);
    return out;
}

std::string trim(std::string const& line, std::string const& whitespace)
{
    std::size_t const b = line.find_first_not_of(whitespace);
    if (b == std::string::npos) return "";
    std::size_t const e = line.find_last_not_of(whitespace);
    return line.substr(b, e - b + 1);
}

std::string reduce(std::string const& line, std::string const& whitespace,
                   std::string const& fill)
{
    std::string s = trim(line, whitespace);
    std::size_t b = s.find_first_of(whitespace);
    while (b != std::string::npos)
    {
        std::size_t const e = s.find_first_not_of(whitespace, b);
        s.replace(b, e - b, fill);
        b = s.find_first_of(whitespace, b + fill.size());
    }
    return s;
}

std::string pad(std::string const& input, std::size_t num, char fill,
                bool right)
{
    if (input.size() >= num) return input;
    std::string const f(num - input.size(), fill);
    return right ? input + f : f + input;
}

std::string strpr(const char* format, ...)
{
    va_list args;
    va_start(args, format);
    va_list copy;
    va_copy(copy, args);
    int const n = std::vsnprintf(nullptr, 0, format, copy);
    va_end(copy);
    std::string out;
    if (n > 0)
    {
        out.resize((std::size_t)n + 1);
        std::vsnprintf(&out[0], out.size(), format, args);
        out.resize((std::size_t)n);
    }
    va_end(args);
    return out;
}

std::string cap(std::string word)
{
    for (auto& c : word) c = (char)std::toupper((unsigned char)c);
    return word;
}

// The column-file header layout of the reference's output files: an
// 80-character rule, the titles, the column table (number, name,
// description), a rule as wide as the data (at least 80), the column
// numbers and names right-aligned over their columns (the first column
// two characters narrower for the comment character and its space), and a
// closing rule as wide as the data.
std::vector<std::string> createFileHeader(
    std::vector<std::string> const& title,
    std::vector<std::size_t> const& colSize,
    std::vector<std::string> const& colName,
    std::vector<std::string> const& colInfo, char const& commentChar)
{
    if (colSize.size() != colName.size() || colName.size() != colInfo.size())
        throw std::runtime_error("ERROR: Could not create file header, "
                                 "column sizes are not equal.\n");
    std::string const rule = std::string(80, commentChar) + "\n";
    std::string const lead(1, commentChar);
    std::vector<std::string> h;
    h.push_back(rule);
    for (auto const& t : title) h.push_back(lead + " " + t + "\n");
    h.push_back(rule);
    std::size_t wName = 0;
    for (auto const& n : colName) wName = std::max(wName, n.size());
    {
        std::ostringstream s;
        s << lead << ' ' << std::left << std::setw(4) << "Col" << ' '
          << std::setw((int)wName) << "Name" << ' ' << "Description\n";
        h.push_back(s.str());
    }
    h.push_back(rule);
    for (std::size_t i = 0; i < colName.size(); ++i)
    {
        std::ostringstream s;
        s << lead << ' ' << std::left << std::setw(4) << i + 1 << ' '
          << std::setw((int)wName) << colName[i] << ' ' << colInfo[i] << '\n';
        h.push_back(s.str());
    }
    std::size_t wData = 0;
    for (auto c : colSize) wData += c + 1;
    if (wData > 0) wData -= 1;
    h.push_back(wData > 80 ? std::string(wData, commentChar) + "\n" : rule);
    std::ostringstream nums, names;
    nums << lead;
    names << lead;
    for (std::size_t i = 0; i < colSize.size(); ++i)
    {
        std::size_t const w = colSize[i] - (i == 0 ? 2 : 0);
        nums << ' ' << std::right << std::setw((int)w) << i + 1;
        std::string const nm = colName[i].substr(0, w);
        names << ' ' << std::right << std::setw((int)w) << nm;
    }
    h.push_back(nums.str() + "\n");
    h.push_back(names.str() + "\n");
    h.push_back(std::string(wData, commentChar) + "\n");
    return h;
}

void appendLinesToFile(std::ofstream& file,
                       std::vector<std::string> const lines)
{
    for (auto const& l : lines) file << l;
}

void appendLinesToFile(FILE* const& file, std::vector<std::string> const lines)
{
    for (auto const& l : lines) std::fputs(l.c_str(), file);
}

double pow_int(double x, int n)
{
    // Exponentiation by squaring (n >= 0; negative n gives 1/x^-n).
    if (n < 0) return 1.0 / pow_int(x, -n);
    double r = 1.0;
    while (n > 0)
    {
        if (n & 1) r *= x;
        x *= x;
        n >>= 1;
    }
    return r;
}

// ----------------------------------------------------------------- Log --

Log& Log::operator<<(std::string const& entry)
{
    addLogEntry(entry);
    return *this;
}

Log& Log::operator<<(std::vector<std::string> const& entries)
{
    addMultipleLogEntries(entries);
    return *this;
}

void Log::addLogEntry(std::string const& entry)
{
    if (silent) return;
    memory.push_back(entry);
    if (writeToStdout)
    {
        std::cout << entry;
        std::cout.flush();
    }
    for (FILE** fp : cFilePointers)
        if (*fp)
        {
            std::fputs(entry.c_str(), *fp);
            std::fflush(*fp);
        }
    for (std::ofstream* s : streamPointers)
        if (s->is_open())
        {
            (*s) << entry;
            s->flush();
        }
}

void Log::addMultipleLogEntries(std::vector<std::string> const& entries)
{
    for (auto const& e : entries) addLogEntry(e);
}

void Log::registerCFilePointer(FILE** const& filePointer)
{
    cFilePointers.push_back(filePointer);
}

void Log::registerStreamPointer(std::ofstream* const& streamPointer)
{
    streamPointers.push_back(streamPointer);
}

// ---------------------------------------------------------- ElementMap --

namespace
{
// Element symbols in order of atomic number (Z = position + 1).
char const* const kSymbols[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg",
    "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr",
    "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
    "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd",
    "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd",
    "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
    "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po",
    "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Np", "Pu", "Am", "Cm",
    "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs",
    "Mt", "Ds", "Rg", "Cn", "Nh", "Fl", "Mc", "Lv", "Ts", "Og"};
std::size_t const kNumSymbols = sizeof(kSymbols) / sizeof(kSymbols[0]);
}  // namespace

std::size_t ElementMap::atomicNumber(std::string const& symbol)
{
    for (std::size_t z = 0; z < kNumSymbols; ++z)
        if (symbol == kSymbols[z]) return z + 1;
    throw std::runtime_error("ERROR: Element \"" + symbol + "\" unknown.\n");
}

std::string ElementMap::symbolFromAtomicNumber(std::size_t const atomicNumber)
{
    if (atomicNumber < 1 || atomicNumber > kNumSymbols)
        throw std::runtime_error("ERROR: Atomic number out of range.\n");
    return kSymbols[atomicNumber - 1];
}

std::size_t ElementMap::registerElements(std::string const& elementLine)
{
    std::vector<std::string> e = split(reduce(elementLine));
    std::sort(e.begin(), e.end(),
              [](std::string const& a, std::string const& b) {
                  return atomicNumber(a) < atomicNumber(b);
              });
    if (!forward.empty())
        throw std::runtime_error("ERROR: Element map not empty.\n");
    forward = e;
    reverse.clear();
    for (std::size_t i = 0; i < forward.size(); ++i) reverse[forward[i]] = i;
    return forward.size();
}

void ElementMap::deregisterElements()
{
    forward.clear();
    reverse.clear();
}

std::size_t ElementMap::index(std::string const& symbol) const
{
    auto it = reverse.find(symbol);
    if (it == reverse.end())
        throw std::out_of_range("ERROR: Element \"" + symbol
                                + "\" not in the element map.\n");
    return it->second;
}

std::string ElementMap::symbol(std::size_t const index) const
{
    if (index >= forward.size())
        throw std::out_of_range("ERROR: Element index out of range.\n");
    return forward[index];
}

std::string ElementMap::getElementsString() const
{
    std::string s;
    for (std::size_t i = 0; i < forward.size(); ++i)
        s += (i ? " " : "") + forward[i];
    return s;
}

std::vector<std::string> ElementMap::info() const
{
    std::vector<std::string> v;
    v.push_back(strpr("********************************\n"));
    v.push_back(strpr("ELEMENTMAP                      \n"));
    v.push_back(strpr("********************************\n"));
    for (std::size_t i = 0; i < forward.size(); ++i)
        v.push_back(strpr("%29zu  : %2s (%3zu)\n", i, forward[i].c_str(),
                          atomicNumber(forward[i])));
    v.push_back(strpr("********************************\n"));
    return v;
}

// ---------------------------------------------------------------- Atom --

void Atom::toNormalizedUnits(double convEnergy, double convLength,
                             double convCharge)
{
    energy *= convEnergy;
    r *= convLength;
    f *= convEnergy / convLength;
    fRef *= convEnergy / convLength;
    charge *= convCharge;
    chargeRef *= convCharge;
    if (hasSymmetryFunctionDerivatives)
        for (auto& v : dEdG) v *= convEnergy;
    for (auto& n : neighbors)
    {
        n.d *= convLength;
        n.dr *= convLength;
    }
}

void Atom::toPhysicalUnits(double convEnergy, double convLength,
                           double convCharge)
{
    energy /= convEnergy;
    r /= convLength;
    f /= convEnergy / convLength;
    fRef /= convEnergy / convLength;
    charge /= convCharge;
    chargeRef /= convCharge;
    if (hasSymmetryFunctionDerivatives)
        for (auto& v : dEdG) v /= convEnergy;
    for (auto& n : neighbors)
    {
        n.d /= convLength;
        n.dr /= convLength;
    }
}

void Atom::free(bool all)
{
    if (all)
    {
        G.clear();
        hasSymmetryFunctions = false;
    }
    dEdG.clear();
    hasSymmetryFunctionDerivatives = false;
}

std::vector<std::string> Atom::info() const
{
    std::vector<std::string> v;
    v.push_back(strpr("index                  : %zu\n", index));
    v.push_back(strpr("indexStructure         : %zu\n", indexStructure));
    v.push_back(strpr("element                : %zu\n", element));
    v.push_back(strpr("numNeighbors           : %zu\n", numNeighbors));
    v.push_back(strpr("numSymmetryFunctions   : %zu\n", numSymmetryFunctions));
    v.push_back(strpr("energy                 : %16.8E\n", energy));
    v.push_back(strpr("r                      : %16.8E %16.8E %16.8E\n", r[0],
                      r[1], r[2]));
    v.push_back(strpr("f                      : %16.8E %16.8E %16.8E\n", f[0],
                      f[1], f[2]));
    v.push_back(strpr("fRef                   : %16.8E %16.8E %16.8E\n",
                      fRef[0], fRef[1], fRef[2]));
    return v;
}

}  // namespace nnp
