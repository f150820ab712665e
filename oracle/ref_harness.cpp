// Reference CPU harness: test infrastructure only, never part of the product.
//
// Runs the reference's own C++ Fock path on one molecule: the basis is
// assembled with CBasisFunction::normalize / CAtomBasis / CMolecularBasis as
// src/pymodule/molecularbasis.py:127-170,289-311 does, the density is placed
// into make_matrix(basis, mat_t::symmetric) as the set_values binding of
// src/python/ExportMath.cpp:258-276 does, then CT4CScreener::partition
// (src/t4c_func/T4CScreener.cpp:63-118) and CFockDriver::compute
// (src/t4c_func/FockDriver.cpp:120-171), timed.  The reference sources are
// compiled where they lie by oracle/ref.mk; nothing of them is copied here.
//
// usage: vlx_ref_fock INPUT OUTPUT
//   INPUT (little endian): int32 natoms; natoms x {int32 Z; double x, y, z (bohr)};
//     natoms x {int32 nshells; nshells x {int32 l, nprim; double exps[nprim], coefs[nprim]}};
//     int32 nao; double D[nao * nao] (row-major, reference AO order);
//     int32 nlabel; char label[nlabel]; double exchange_factor, omega; int32 ithreshold, repeats
//   OUTPUT: double G[nao * nao] of the last repeat; timings as one JSON line on stdout.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

#include "AtomBasis.hpp"
#include "BasisFunction.hpp"
#include "FockDriver.hpp"
#include "Matrix.hpp"
#include "MatrixFunc.hpp"
#include "MolecularBasis.hpp"
#include "Molecule.hpp"
#include "Point.hpp"
#include "T4CScreener.hpp"

namespace {

struct Reader {
    FILE* f;
    template <typename T>
    T get()
    {
        T v{};
        if (std::fread(&v, sizeof(T), 1, f) != 1) throw std::runtime_error("short input");
        return v;
    }
    template <typename T>
    std::vector<T> vec(size_t n)
    {
        std::vector<T> v(n);
        if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("short input");
        return v;
    }
};

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s INPUT OUTPUT\n", argv[0]);
        return 2;
    }
    try {
        FILE* fin = std::fopen(argv[1], "rb");
        if (!fin) throw std::runtime_error("cannot open input");
        Reader rd{fin};
        const int natoms = rd.get<int32_t>();
        std::vector<int> ids(natoms);
        std::vector<TPoint<double>> xyz;
        for (int a = 0; a < natoms; a++) {
            ids[a] = rd.get<int32_t>();
            const double x = rd.get<double>(), y = rd.get<double>(), z = rd.get<double>();
            xyz.emplace_back(std::array<double, 3>{x, y, z});
        }
        CMolecule molecule(ids, xyz, "au");
        CMolecularBasis basis;
        for (int a = 0; a < natoms; a++) {
            CAtomBasis atom_basis;
            const int nshells = rd.get<int32_t>();
            for (int s = 0; s < nshells; s++) {
                const int l = rd.get<int32_t>(), np = rd.get<int32_t>();
                const auto exps = rd.vec<double>(np);
                const auto coefs = rd.vec<double>(np);
                CBasisFunction bf(exps, coefs, l);
                bf.normalize();
                atom_basis.add(bf);
            }
            atom_basis.set_identifier(ids[a]);
            atom_basis.set_name("HARNESS");
            basis.add(atom_basis);
        }
        const int nao = rd.get<int32_t>();
        const auto dens = rd.vec<double>((size_t)nao * nao);
        const int nlabel = rd.get<int32_t>();
        const auto lab = rd.vec<char>(nlabel);
        const std::string label(lab.begin(), lab.end());
        const double xf = rd.get<double>(), omega = rd.get<double>();
        const int ithreshold = rd.get<int32_t>(), repeats = std::max(1, (int)rd.get<int32_t>());
        std::fclose(fin);

        auto dmat = matfunc::make_matrix(basis, mat_t::symmetric);
        for (const auto& key : dmat.angular_pairs()) {
            auto sub = dmat.sub_matrix(key);
            const auto dims = sub->get_dimensions();
            for (size_t i = 0; i < dims[2]; i++)
                for (size_t j = 0; j < dims[3]; j++) sub->at({i, j}) = dens[(i + dims[0]) * nao + j + dims[1]];
        }

        auto t0 = std::chrono::steady_clock::now();
        CT4CScreener screener;
        screener.partition(basis, molecule, "eri");
        const double t_screen = seconds_since(t0);

        CFockDriver driver;
        CMatrix fock;
        std::vector<double> times;
        for (int r = 0; r < repeats; r++) {
            t0 = std::chrono::steady_clock::now();
            fock = driver.compute(screener, dmat, label, xf, omega, ithreshold);
            times.push_back(seconds_since(t0));
        }
        const auto full = fock.full_matrix();
        if ((int)full.number_of_rows() != nao) throw std::runtime_error("unexpected Fock dimensions");
        FILE* fout = std::fopen(argv[2], "wb");
        if (!fout) throw std::runtime_error("cannot open output");
        std::fwrite(full.data(), sizeof(double), (size_t)nao * nao, fout);
        std::fclose(fout);
        std::sort(times.begin(), times.end());
        std::printf("{\"nao\": %d, \"screener_s\": %.6f, \"median_s\": %.6f, \"min_s\": %.6f, \"repeats\": %d}\n", nao,
                    t_screen, times[times.size() / 2], times.front(), repeats);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "vlx_ref_fock: %s\n", e.what());
        return 1;
    }
    return 0;
}
