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

all: $(OUT)/vlx_ref_fock

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
