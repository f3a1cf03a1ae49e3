# ref.mk - builds the reference CPU path (checker only) from its sources in
# place under /root/reference into oracle/_ref/.
#
#   make -f oracle/ref.mk            # oracle/_ref/3dmhd (periodic gstate dumps)
#
# The reference's own makefile is not used: this recipe compiles the CPU
# translation units directly with MPICH's mpicxx (g++ 11 as the host compiler),
# with the reference's bit-exact host flags (-O2 -ffp-contract=off -fopenmp;
# reference makefile:104, :155-166).  Outputs go only to oracle/_ref/, which
# git ignores.  Nothing here is product code: the product never links or runs
# anything built by this file.
REF     ?= /root/reference
HERE    := $(patsubst %/,%,$(dir $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)/_ref
MPICXX  ?= /opt/conda/bin/mpicxx
export MPICH_CXX := g++

CXXFLAGS := -O2 -std=c++17 -ffp-contract=off -fopenmp -I$(REF)/include
SRCS := params topology grid ode ran2 static_ic fluxes bcon communicate step
OBJS := $(addprefix $(OUT)/obj/,$(addsuffix .o,$(SRCS)))

all: $(OUT)/3dmhd $(OUT)/3dmhd_nodump $(OUT)/hmean $(OUT)/stepdt

$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(CXXFLAGS) -c -o $@ $<

$(OUT)/obj/main_dump.o: $(REF)/src/main.cpp
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(CXXFLAGS) -DR3D_DUMP_MODE=2 -c -o $@ $<

$(OUT)/obj/main_nodump.o: $(REF)/src/main.cpp
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(CXXFLAGS) -DR3D_DUMP_MODE=0 -c -o $@ $<

$(OUT)/3dmhd: $(OUT)/obj/main_dump.o $(OBJS)
	$(MPICXX) $(CXXFLAGS) -o $@ $^ -lm

$(OUT)/3dmhd_nodump: $(OUT)/obj/main_nodump.o $(OBJS)
	$(MPICXX) $(CXXFLAGS) -o $@ $^ -lm

.PHONY: all

# the reference's horizontal_mean alone, driven by oracle/ref_hmean.cpp
$(OUT)/obj/ref_hmean.o: $(HERE)/ref_hmean.cpp
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(CXXFLAGS) -c -o $@ $<

$(OUT)/hmean: $(OUT)/obj/ref_hmean.o $(OBJS)
	$(MPICXX) $(CXXFLAGS) -o $@ $^ -lm

# the reference's STEP time-step limits (LMAG included), driven by
# oracle/ref_stepdt.cpp; the link sends STEP's first FLUXES call there
FLUXES_SYM := _ZN3r3d6fluxesERKNS_6ParamsERKNS_7DerivedERKNS_8TopologyEPKdSA_RNS_8SimStateE
$(OUT)/obj/ref_stepdt.o: $(HERE)/ref_stepdt.cpp
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(CXXFLAGS) -c -o $@ $<

$(OUT)/stepdt: $(OUT)/obj/ref_stepdt.o $(OBJS)
	$(MPICXX) $(CXXFLAGS) -Wl,--wrap=$(FLUXES_SYM) -o $@ $^ -lm
