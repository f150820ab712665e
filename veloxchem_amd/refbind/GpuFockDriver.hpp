// Reference-side drop-in of the GPU Fock build: C++ host code that a VeloxChem
// build links next to CFockDriver (src/t4c_func/FockDriver.hpp:36-133).  It
// takes the reference's own inputs (CT4CScreener, CMatrix / CMatrices, label,
// exchange factor, omega, ithreshold) and returns a CMatrix shaped like the
// density, computed by libvlxfock.so on an MI355X.  Compiled against the
// reference headers; the pybind11 export is ExportGpuFock.cpp.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "Matrices.hpp"
#include "Matrix.hpp"
#include "T4CScreener.hpp"

struct vlx_context;
struct vlx_screener;

// Shell table of the library's vlx_screener_create, rebuilt from a reference screener's pair
// blocks (every shell appears in its own diagonal pair).
struct CGpuShellTable {
    std::vector<int> l, nprim, prim_offset, ao_base, ao_stride, atom;
    std::vector<double> exps, norms, centers;
    int nao = 0;
};

// The shell table behind a CT4CScreener; throws std::runtime_error when the pair blocks do not
// describe a complete basis.
auto make_shell_table(const CT4CScreener& screener) -> CGpuShellTable;

class CGpuFockDriver {
   public:
    explicit CGpuFockDriver(int device = 0);
    ~CGpuFockDriver();
    CGpuFockDriver(const CGpuFockDriver&) = delete;
    CGpuFockDriver& operator=(const CGpuFockDriver&) = delete;

    // CFockDriver::set_block_size_factor (src/t4c_func/FockDriver.cpp:356-376): kept for the
    // interface; the GPU schedule does not depend on it.
    auto set_block_size_factor(const int factor) -> void;

    // CFockDriver::compute(screener, density, label, exchange_factor, omega, ithreshold)
    // (src/t4c_func/FockDriver.cpp:120-171): the Fock matrix of one density, of the density's
    // type; symmetric and general densities.
    auto compute(const CT4CScreener& screener, const CMatrix& density, const std::string& label,
                 const double exchange_factor, const double omega, const int ithreshold) const -> CMatrix;

    // The rank/nodes variant (src/t4c_func/FockDriver.cpp:234-289): one GPU computes the whole
    // matrix, so only rank 0 of 1 node is accepted.
    auto compute(const CT4CScreener& screener, const int rank, const int nodes, const CMatrix& density,
                 const std::string& label, const double exchange_factor, const double omega,
                 const int ithreshold) const -> CMatrix;

    // Several densities with one label each (src/t4c_func/FockDriver.cpp:176-229).
    auto compute(const CT4CScreener& screener, const CMatrices& densities, const std::vector<std::string>& labels,
                 const double exchange_factor, const double omega, const int ithreshold) const -> CMatrices;

   private:
    auto _screener_for(const CT4CScreener& screener) const -> vlx_screener*;

    int _device;
    int _block_size_factor = 8;
    vlx_context* _ctx = nullptr;
    // library screeners by the reference screener's address, refreshed when its pair blocks change
    mutable std::map<const CT4CScreener*, std::pair<std::vector<CBlockedGtoPairBlock>, vlx_screener*>> _screeners;
};
