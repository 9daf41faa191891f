# ref.mk -- builds the reference's own hydro2d module, as a CPU checker and
# CPU baseline, from its sources where they lie (nothing is copied).
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Outputs (git-ignored, travel with the tree):
#   oracle/_ref/hydro_ref        -O2 -ffp-contract=off: parity fixtures
#   oracle/_ref/hydro_ref_fast   -O3 -march=x86-64-v3 -fopenmp: CPU baseline
#   oracle/_ref/hydro_ref_gpu    (make ... gpu) the parity driver with the
#                                reference's solvers dispatched to libhydrogpu.so
#                                (the reference builds with -O3 -march=native,
#                                CMakeLists.txt:69-74; x86-64-v3 keeps the
#                                binary runnable on other x86 hosts)
REF ?= /root/reference
OUT := oracle/_ref
SRC := $(REF)/source
REF_SRCS := $(SRC)/common/common.cpp $(SRC)/common/linear_algebra.cpp \
            $(SRC)/control/experiment.cpp $(SRC)/control/module.cpp \
            $(SRC)/control/console.cpp $(SRC)/control/pvar.cpp \
            $(SRC)/control/experiment_ref.cpp $(SRC)/hydro2dmpi/hydro2d.cpp \
            $(SRC)/hydro2dmpi/output.cpp $(SRC)/test.cpp $(REF)/tools/revision.cpp
CXX ?= g++
COMMON := -std=c++11 -w -I$(SRC) -DMODULE_HYDRO_2D -fopenmp -pthread

all: $(OUT)/hydro_ref $(OUT)/hydro_ref_fast $(OUT)/general.hydroconf

# hydro_ref_gpu: the same driver with the reference's pressure (and velocity)
# LinearSolver replaced by the GpuLinearSolver adapter over libhydrogpu.so
# (oracle/ref_gpu_dispatch.cpp); built when the library exists.
GPU_LIB := hydro_amd/libhydrogpu.so
gpu: $(OUT)/hydro_ref_gpu

# The reference binary reads every parameter it uses; its defaults file
# goes next to it (an output of this build, not a committed file).
$(OUT)/general.hydroconf: $(REF)/examples/general.hydroconf
	mkdir -p $(OUT)
	cp $< $@

$(OUT)/hydro_ref: oracle/ref_shim.cpp $(REF_SRCS)
	mkdir -p $(OUT)
	$(CXX) $(COMMON) -O2 -ffp-contract=off -o $@ oracle/ref_shim.cpp $(REF_SRCS)

$(OUT)/hydro_ref_fast: oracle/ref_shim.cpp $(REF_SRCS)
	mkdir -p $(OUT)
	$(CXX) $(COMMON) -O3 -march=x86-64-v3 -o $@ oracle/ref_shim.cpp $(REF_SRCS)

$(OUT)/hydro_ref_gpu: oracle/ref_shim.cpp oracle/ref_gpu_dispatch.cpp $(REF_SRCS) $(GPU_LIB)
	mkdir -p $(OUT)
	$(CXX) $(COMMON) -O2 -ffp-contract=off -Iinclude -o $@ oracle/ref_shim.cpp \
	  oracle/ref_gpu_dispatch.cpp $(REF_SRCS) -Lhydro_amd -lhydrogpu \
	  -Wl,-rpath,'$$ORIGIN/../../hydro_amd'

# hydro_ref3d: the same driver on the reference's 3D module (hydro3d.cpp,
# mesh3d.cpp; MODULE_HYDRO_3D), fixtures of the 3D step
ref3d: $(OUT)/hydro_ref3d
$(OUT)/hydro_ref3d: oracle/ref_shim.cpp $(REF_SRCS)
	mkdir -p $(OUT)
	$(CXX) $(COMMON) -DMODULE_HYDRO_3D -DHR_DIM=3 -O2 -ffp-contract=off -o $@ \
	  oracle/ref_shim.cpp $(REF_SRCS) $(SRC)/hydro2dmpi/hydro3d.cpp $(SRC)/hydro2dmpi/mesh3d.cpp

.PHONY: all gpu ref3d
