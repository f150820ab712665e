# Reference CPU harness (test infrastructure only): compiles the reference's
# Fock path where it lies (REF) and links it with oracle/ref_harness.cpp into
# oracle/_ref/vlx_ref_fock.  Nothing of the reference is copied into the
# repository; objects and the binary stay in oracle/_ref/ (git-ignored).
#   make -f oracle/ref.mk -j8          (from the repository root)
REF ?= /root/reference/src
OUT := oracle/_ref
CXX ?= g++
# the reference's release flags (-O3, OpenMP), with a portable x86-64-v3 target
# in place of -march=native so the binary runs on other hosts
CXXFLAGS := -std=c++20 -O3 -march=x86-64-v3 -fopenmp -w
DIRS := general math moldata orbdata t2c_func t4c_func t4c_electron_repulsion
INC := $(foreach d,$(DIRS),-I$(REF)/$(d))
# sources that need LAPACK/Eigen (and the two that call them) are not on the Fock path
SKIP := $(REF)/math/DenseDiagonalizer.cpp $(REF)/math/DenseLinearAlgebra.cpp \
        $(REF)/moldata/DispersionModel.cpp $(REF)/moldata/PartialCharges.cpp
SRCS := $(filter-out $(SKIP),$(foreach d,$(DIRS),$(wildcard $(REF)/$(d)/*.cpp)))
OBJS := $(patsubst $(REF)/%.cpp,$(OUT)/obj/%.o,$(SRCS))

# the reference-side drop-in (veloxchem_amd/refbind): CGpuFockDriver against the reference headers,
# its pybind11 export (compiled, not linked: the reference's Python module registers the argument
# types) and a harness that runs it beside CFockDriver on the same CT4CScreener and CMatrix
BIND := veloxchem_amd/refbind
PYBIND_INC := $(shell python3 -c "import pybind11; print(pybind11.get_include())")
PY_INC := $(shell python3 -c "import sysconfig; print(sysconfig.get_paths()['include'])")

all: $(OUT)/vlx_ref_fock $(OUT)/vlx_ref_gpu_fock $(OUT)/refbind/ExportGpuFock.o

$(OUT)/obj/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c -o $@ $<

$(OUT)/harness.o: oracle/ref_harness.cpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) $(INC) -c -o $@ $<

$(OUT)/vlx_ref_fock: $(OUT)/harness.o $(OBJS)
	$(CXX) -fopenmp -o $@ $^
	strip $@

.PHONY: all

$(OUT)/refbind/GpuFockDriver.o: $(BIND)/GpuFockDriver.cpp $(BIND)/GpuFockDriver.hpp include/vlxfock.h
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c -o $@ $<

$(OUT)/refbind/ExportGpuFock.o: $(BIND)/ExportGpuFock.cpp $(BIND)/GpuFockDriver.hpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -fPIC $(INC) -I$(REF)/python -I$(PYBIND_INC) -I$(PY_INC) -c -o $@ $<

$(OUT)/ref_gpu_harness.o: oracle/ref_gpu_harness.cpp $(BIND)/GpuFockDriver.hpp
	@mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) $(INC) -c -o $@ $<

# libvlxfock.so is found next to the package at run time ($ORIGIN = oracle/_ref)
$(OUT)/vlx_ref_gpu_fock: $(OUT)/ref_gpu_harness.o $(OUT)/refbind/GpuFockDriver.o $(OBJS) veloxchem_amd/libvlxfock.so
	$(CXX) -fopenmp -o $@ $(OUT)/ref_gpu_harness.o $(OUT)/refbind/GpuFockDriver.o $(OBJS) \
	    -Lveloxchem_amd -lvlxfock -Wl,-rpath,'$$ORIGIN/../../veloxchem_amd'
	strip $@
