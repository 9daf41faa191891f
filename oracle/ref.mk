# ref.mk -- builds the reference's own hydro2d module, as a CPU checker and
# CPU baseline, from its sources where they lie (nothing is copied).
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Outputs (git-ignored, travel with the tree):
#   oracle/_ref/hydro_ref        -O2 -ffp-contract=off: parity fixtures
#   oracle/_ref/hydro_ref_fast   -O3 -march=x86-64-v3 -fopenmp: CPU baseline
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

.PHONY: all
