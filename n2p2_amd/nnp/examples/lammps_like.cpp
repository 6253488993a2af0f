// Drives nnp::InterfaceLammps the way pair_style hdnnp does
// (pair_hdnnp.cpp:81-114, 226-247): local atoms with LAMMPS types, a full
// neighbour list over local and ghost atoms (dx = x_i - x_j), process(),
// then the energy and the forces of all local and ghost atoms.
//
// Usage: lammps_like <nnp directory> <domain file>
// Domain file: "ntypes nlocal nall cutoff", then nall lines "type x y z",
// local atoms first.  Prints "energy E" and "force k fx fy fz" for every
// atom k (17 significant digits, LAMMPS units = the model's physical units).
#include "InterfaceLammps.h"

#include <cstdio>
#include <exception>
#include <fstream>
#include <vector>

int main(int argc, char* argv[])
{
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: %s <nnp directory> <domain file>\n",
                     argv[0]);
        return 1;
    }
    try
    {
        std::ifstream in(argv[2]);
        int ntypes = 0, nlocal = 0, nall = 0;
        double cutoff = 0.0;
        in >> ntypes >> nlocal >> nall >> cutoff;
        std::vector<int> type(nall);
        std::vector<double> x(3 * (size_t)nall);
        for (int k = 0; k < nall; ++k)
            in >> type[k] >> x[3 * k] >> x[3 * k + 1] >> x[3 * k + 2];
        if (!in) throw std::runtime_error("ERROR: Bad domain file.\n");

        nnp::InterfaceLammps lmp;
        lmp.initialize(argv[1], "", false, false, 0, -1, 1.0, 1.0, cutoff,
                       ntypes, 0);
        std::vector<int64_t> tag(nall);
        for (int k = 0; k < nall; ++k) tag[k] = k + 1;
        lmp.setLocalAtoms(nlocal, type.data());
        lmp.setLocalTags(tag.data());
        double const rc = lmp.getMaxCutoffRadius();
        double const rc2 = rc * rc;
        for (int i = 0; i < nlocal; ++i)
            for (int j = 0; j < nall; ++j)
            {
                double const dx = x[3 * i] - x[3 * j];
                double const dy = x[3 * i + 1] - x[3 * j + 1];
                double const dz = x[3 * i + 2] - x[3 * j + 2];
                double const d2 = dx * dx + dy * dy + dz * dz;
                if (j != i && d2 <= rc2)
                    lmp.addNeighbor(i, j, tag[j], type[j], dx, dy, dz, d2);
            }
        lmp.finalizeNeighborList();
        lmp.process();
        std::vector<double> f(3 * (size_t)nall, 0.0);
        std::vector<double*> rows(nall);
        for (int k = 0; k < nall; ++k) rows[k] = &f[3 * k];
        double* const* const fp = rows.data();
        lmp.getForces(fp);
        std::printf("energy %.17g\n", lmp.getEnergy());
        for (int k = 0; k < nall; ++k)
            std::printf("force %d %.17g %.17g %.17g\n", k, f[3 * k],
                        f[3 * k + 1], f[3 * k + 2]);
    }
    catch (std::exception const& e)
    {
        std::fprintf(stderr, "%s", e.what());
        return 2;
    }
    return 0;
}
