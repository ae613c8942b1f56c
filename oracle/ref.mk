# Builds the reference's own hot path (derekeverett/cpu-vh) into
# oracle/_ref/libcpuvh_ref_<variant>.so, straight from the sources under
# $(REF).  Test infrastructure only: the product never links these.
#
#   conf       default header configuration (shear only, conformal EoS)
#   conf_bulk  -DPI (bulk viscous pressure on), conformal EoS
#   lat        HotQCD+HRG table EoS, shear only
#   lat_bulk   HotQCD+HRG table EoS, bulk on
#
# The EoS is selected by a #define in rhic/include/EquationOfState.h:13-15
# (CONFORMAL_EOS is defined unconditionally), so the table variants read a
# configured copy of that one header, generated below into $(CFG) outside the
# repository; every .cpp is still compiled from where it lies, with gcc's -I-
# so that "../include/EquationOfState.h" resolves to the configured header.
#
# Golden runs: -ffp-contract=off (no FMA contraction) and OMP_NUM_THREADS=1
# (the ghost-cell loops share their index variables across threads,
# rhic/src/DynamicalVariables.cpp:290,324,358).

REF ?= /root/reference
OUT ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CFG ?= /tmp/cpuvh_ref_cfg
CXX ?= g++
CXXFLAGS_REF = -O3 -fopenmp -std=c++11 -ffp-contract=off -march=x86-64-v3 -fPIC -fvisibility=hidden -w
HERE := $(dir $(lastword $(MAKEFILE_LIST)))

SRCS = FullyDiscreteKurganovTadmorScheme SemiDiscreteKurganovTadmorScheme HalfSiteExtrapolation \
       FluxLimiter LocalPropagationSpeed SpectralRadius FluxFunctions SourceTerms \
       EnergyMomentumTensor EquationOfState DynamicalVariables Vorticity
REF_CPP = $(addprefix $(REF)/rhic/src/,$(addsuffix .cpp,$(SRCS)))
HARNESS = $(HERE)ref_harness.cpp $(HERE)ref_fo_unit.cpp

INC_CONF = -I$(REF)/rhic/include -I$(REF)/freezeout -I$(REF)/freezeout/cornelius-c++-1.3
INC_LAT  = -I- -I$(CFG)/src -I$(CFG)/fo -I$(CFG)/include -I$(REF)/freezeout -I$(REF)/freezeout/cornelius-c++-1.3

all: $(OUT)/libcpuvh_ref_conf.so $(OUT)/libcpuvh_ref_conf_bulk.so $(OUT)/libcpuvh_ref_lat.so $(OUT)/libcpuvh_ref_lat_bulk.so

$(OUT):
	mkdir -p $(OUT)

$(CFG)/include/EquationOfState.h: $(REF)/rhic/include/EquationOfState.h
	mkdir -p $(CFG)/include $(CFG)/src $(CFG)/fo $(CFG)/rhic
	for h in $(REF)/rhic/include/*.h; do ln -sf $$h $(CFG)/include/; done
	rm -f $(CFG)/include/EquationOfState.h
	sed -e 's|^#define CONFORMAL_EOS|//#define CONFORMAL_EOS|' -e 's|^//#define HOTQCDHRG_EOS|#define HOTQCDHRG_EOS|' $< > $@
	ln -sfn $(CFG)/include $(CFG)/rhic/include

$(OUT)/libcpuvh_ref_conf.so: $(REF_CPP) $(HARNESS) | $(OUT)
	$(CXX) $(CXXFLAGS_REF) $(INC_CONF) -shared -o $@ $(REF_CPP) $(HARNESS) -Wl,-Bsymbolic -Wl,--no-undefined

$(OUT)/libcpuvh_ref_conf_bulk.so: $(REF_CPP) $(HARNESS) | $(OUT)
	$(CXX) $(CXXFLAGS_REF) -DPI $(INC_CONF) -shared -o $@ $(REF_CPP) $(HARNESS) -Wl,-Bsymbolic -Wl,--no-undefined

$(OUT)/libcpuvh_ref_lat.so: $(REF_CPP) $(HARNESS) $(CFG)/include/EquationOfState.h | $(OUT)
	$(CXX) $(CXXFLAGS_REF) $(INC_LAT) -shared -o $@ $(REF_CPP) $(HARNESS) -Wl,-Bsymbolic -Wl,--no-undefined

$(OUT)/libcpuvh_ref_lat_bulk.so: $(REF_CPP) $(HARNESS) $(CFG)/include/EquationOfState.h | $(OUT)
	$(CXX) $(CXXFLAGS_REF) -DPI $(INC_LAT) -shared -o $@ $(REF_CPP) $(HARNESS) -Wl,-Bsymbolic -Wl,--no-undefined

.PHONY: all
