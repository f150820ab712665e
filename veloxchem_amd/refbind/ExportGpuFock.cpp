// pybind11 export of the GPU Fock build for the reference's Python module:
// `_GpuFockDriver` has the methods of `_FockDriver`
// (src/python/ExportT4CIntegrals.cpp:48-124) that the Fock path uses, with the
// same argument lists, so that src/pymodule/fockdriver.py can hold either.  A
// VeloxChem build adds this file and GpuFockDriver.cpp to its sources, links
// libvlxfock.so and calls export_gpu_fock(m) next to export_t4cintegrals(m);
// the argument types (T4CScreener, Matrix, Matrices) are the module's own
// registered classes.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "ExportHelpers.hpp"
#include "GpuFockDriver.hpp"

namespace py = pybind11;
using namespace py::literals;

void export_gpu_fock(py::module& m)
{
    PyClass<CGpuFockDriver>(m, "_GpuFockDriver")
        .def(py::init<int>(), "device"_a = 0)
        .def("_set_block_size_factor", &CGpuFockDriver::set_block_size_factor, "Sets block size factor.")
        .def(
            "_compute_fock_omp",
            [](const CGpuFockDriver& drv, const CT4CScreener& screener, const CMatrix& density, const std::string& label,
               const double exchange_factor, const double omega, const int ithreshold) -> CMatrix {
                py::gil_scoped_release release;
                return drv.compute(screener, density, label, exchange_factor, omega, ithreshold);
            },
            "Computes single Fock matrix of requested type for two-electron integrals screener.")
        .def(
            "_compute_fock_omp",
            [](const CGpuFockDriver& drv, const CT4CScreener& screener, const CMatrices& densities,
               const std::vector<std::string>& labels, const double exchange_factor, const double omega,
               const int ithreshold) -> CMatrices {
                py::gil_scoped_release release;
                return drv.compute(screener, densities, labels, exchange_factor, omega, ithreshold);
            },
            "Computes Fock matrices of requested type for two-electron integrals screener.")
        .def(
            "_compute_local_fock",
            [](const CGpuFockDriver& drv, const CT4CScreener& screener, const int rank, const int nodes,
               const CMatrix& density, const std::string& label, const double exchange_factor, const double omega,
               const int ithreshold) -> CMatrix {
                py::gil_scoped_release release;
                return drv.compute(screener, rank, nodes, density, label, exchange_factor, omega, ithreshold);
            },
            "Computes single Fock matrix of requested type for two-electron integrals screener.")
        .def(
            "_compute_local_fock",
            [](const CGpuFockDriver& drv, const CT4CScreener& screener, const int rank, const int nodes,
               const CMatrices& densities, const std::vector<std::string>& labels, const double exchange_factor,
               const double omega, const int ithreshold) -> CMatrices {
                if (rank != 0 || nodes != 1)
                    throw std::runtime_error("_GpuFockDriver: one GPU builds the whole matrix (rank 0 of 1 node only)");
                py::gil_scoped_release release;
                return drv.compute(screener, densities, labels, exchange_factor, omega, ithreshold);
            },
            "Computes Fock matrices of requested type for two-electron integrals screener.");
}
