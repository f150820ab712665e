// Reference-side binding check: test infrastructure only, never part of the product.
//
// Builds one molecule's CMolecularBasis, CMolecule, CT4CScreener and density
// CMatrix exactly as vlx_ref_fock does (same input format, see
// oracle/ref_harness.cpp), then runs the reference's CFockDriver::compute
// (src/t4c_func/FockDriver.cpp:120-171 and the CMatrices overload :176-229)
// and the drop-in CGpuFockDriver (veloxchem_amd/refbind) on the same objects:
// one density with the input's label, and a two-density CMatrices call with
// labels {label, VLX_HARNESS_LABEL2 (default "k")} (the second density is half
// the first).  The density is
// stored as mat_t::symmetric, or as mat_t::general when the environment sets
// VLX_HARNESS_GENERAL=1 (then the input density need not be symmetric).
//
// usage: vlx_ref_gpu_fock INPUT OUTPUT
//   OUTPUT: G_reference, G_gpu (nao x nao doubles each); one JSON line on stdout
//   with the largest elementwise differences.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "AtomBasis.hpp"
#include "BasisFunction.hpp"
#include "FockDriver.hpp"
#include "Matrices.hpp"
#include "Matrix.hpp"
#include "MatrixFunc.hpp"
#include "MolecularBasis.hpp"
#include "Molecule.hpp"
#include "T4CScreener.hpp"
#include "../veloxchem_amd/refbind/GpuFockDriver.hpp"

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

double max_diff(const CMatrix& a, const CMatrix& b)
{
    const auto fa = a.full_matrix(), fb = b.full_matrix();
    const size_t n = fa.number_of_rows() * fa.number_of_columns();
    if (fb.number_of_rows() * fb.number_of_columns() != n) return INFINITY;
    double m = 0.0;
    for (size_t i = 0; i < n; i++) m = std::max(m, std::fabs(fa.data()[i] - fb.data()[i]));
    return m;
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
        const int ithreshold = rd.get<int32_t>();
        (void)rd.get<int32_t>();  // repeats (unused)
        std::fclose(fin);

        const char* gen = std::getenv("VLX_HARNESS_GENERAL");
        const mat_t type = (gen && std::string(gen) == "1") ? mat_t::general : mat_t::symmetric;
        auto dmat = matfunc::make_matrix(basis, type);
        for (const auto& key : dmat.angular_pairs()) {
            auto sub = dmat.sub_matrix(key);
            const auto dims = sub->get_dimensions();
            for (size_t i = 0; i < dims[2]; i++)
                for (size_t j = 0; j < dims[3]; j++) sub->at({i, j}) = dens[(i + dims[0]) * nao + j + dims[1]];
        }

        CT4CScreener screener;
        screener.partition(basis, molecule, "eri");

        CFockDriver ref_drv;
        CGpuFockDriver gpu_drv(0);
        const CMatrix f_ref = ref_drv.compute(screener, dmat, label, xf, omega, ithreshold);
        const CMatrix f_gpu = gpu_drv.compute(screener, dmat, label, xf, omega, ithreshold);
        // the rank/nodes overload of the Python layer's _compute_local_fock
        const CMatrix f_loc = gpu_drv.compute(screener, 0, 1, dmat, label, xf, omega, ithreshold);

        auto half = dmat;
        half.scale(0.5);
        CMatrices dens2;
        dens2.add(dmat, "0");
        dens2.add(half, "1");
        const char* l2 = std::getenv("VLX_HARNESS_LABEL2");
        const std::vector<std::string> labels{label, l2 ? std::string(l2) : std::string("k")};
        const CMatrices m_ref = ref_drv.compute(screener, dens2, labels, xf, omega, ithreshold);
        const CMatrices m_gpu = gpu_drv.compute(screener, dens2, labels, xf, omega, ithreshold);
        double dm = 0.0;
        for (const auto& key : m_ref.keys()) dm = std::max(dm, max_diff(*m_ref.matrix(key), *m_gpu.matrix(key)));

        const auto full_ref = f_ref.full_matrix(), full_gpu = f_gpu.full_matrix();
        if ((int)full_ref.number_of_rows() != nao) throw std::runtime_error("unexpected Fock dimensions");
        FILE* fout = std::fopen(argv[2], "wb");
        if (!fout) throw std::runtime_error("cannot open output");
        std::fwrite(full_ref.data(), sizeof(double), (size_t)nao * nao, fout);
        std::fwrite(full_gpu.data(), sizeof(double), (size_t)nao * nao, fout);
        std::fclose(fout);
        std::printf("{\"nao\": %d, \"general\": %d, \"same_type\": %d, \"max_abs_diff\": %.3e, "
                    "\"max_abs_diff_local\": %.3e, \"max_abs_diff_matrices\": %.3e}\n",
                    nao, type == mat_t::general ? 1 : 0, f_gpu.get_type() == f_ref.get_type() ? 1 : 0,
                    max_diff(f_ref, f_gpu), max_diff(f_ref, f_loc), dm);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "vlx_ref_gpu_fock: %s\n", e.what());
        return 1;
    }
    return 0;
}
