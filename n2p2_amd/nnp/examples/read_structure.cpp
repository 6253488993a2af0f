// Host-only check of the nnp layer's input.data reader: reads the first
// structure of a file with the elements of a settings file and prints the
// atom count, the cell and every position after Structure::remap as
// hexadecimal floats (exact), one "atom i element x y z" line per atom.
//
// Usage: read_structure <input.nn> <input.data>
#include "ElementMap.h"
#include "Structure.h"
#include "utility.h"

#include <cstdio>
#include <exception>
#include <fstream>
#include <string>

int main(int argc, char* argv[])
{
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: %s <input.nn> <input.data>\n", argv[0]);
        return 1;
    }
    try
    {
        std::ifstream settings(argv[1]);
        std::string line, elements;
        while (std::getline(settings, line))
        {
            std::size_t const hash = line.find('#');
            if (hash != std::string::npos) line.erase(hash);
            std::vector<std::string> w = nnp::split(nnp::reduce(line));
            if (!w.empty() && w[0] == "elements")
                elements = nnp::reduce(line).substr(9);
        }
        nnp::ElementMap map;
        map.registerElements(elements);
        nnp::Structure s;
        s.setElementMap(map);
        s.readFromFile(std::string(argv[2]));
        std::printf("atoms %zu periodic %d\n", s.numAtoms, (int)s.isPeriodic);
        for (int k = 0; k < 3; ++k)
            std::printf("box %a %a %a\n", s.box[k][0], s.box[k][1],
                        s.box[k][2]);
        for (auto const& a : s.atoms)
            std::printf("atom %zu %zu %a %a %a\n", a.index, a.element, a.r[0],
                        a.r[1], a.r[2]);
    }
    catch (std::exception const& e)
    {
        std::fprintf(stderr, "%s", e.what());
        return 2;
    }
    return 0;
}
