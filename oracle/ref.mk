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

# The reference's own driver (src/main.cpp) with its --backend gpu loop
# (src/main.cpp:234-275), linked with the INTEGRATION.md shim and the product
# library: the reference calls the library through its include/gpu.hpp
# surface.  main.cpp is compiled with R3D_HAVE_GPU and cudaSetDevice defined
# away (the library opens cfg.device itself); no CUDA is involved.  The shim
# source is the code block of INTEGRATION.md, extracted into _ref/dropin/.
ROOT   := $(abspath $(HERE)/..)
DROPIN := $(OUT)/dropin
all: $(OUT)/3dmhd_gpu

$(DROPIN)/gpu_hip.cpp: $(ROOT)/INTEGRATION.md
	@mkdir -p $(DROPIN)
	python3 -c "import re,sys; b=re.findall(r'\`\`\`(?:cpp|c\+\+)\n(.*?)\`\`\`', open(sys.argv[1]).read(), re.S); assert len(b)==1; open(sys.argv[2],'w').write(b[0])" $< $@

$(OUT)/obj/gpu_hip.o: $(DROPIN)/gpu_hip.cpp $(ROOT)/include/r3d_hip.h $(ROOT)/include/r3d_gpu.hpp
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(CXXFLAGS) -I$(ROOT)/include -c -o $@ $<

$(OUT)/obj/main_gpu.o: $(REF)/src/main.cpp
	@mkdir -p $(OUT)/obj
	$(MPICXX) $(CXXFLAGS) -DR3D_DUMP_MODE=2 -DR3D_HAVE_GPU '-DcudaSetDevice(d)=((void)(d))' -c -o $@ $<

# Linked by g++ with MPICH's libraries named by path: the ROCm runtime needs
# the system libstdc++, which the wrapper's -L/rpath of its own lib directory
# (an older libstdc++) would hide at link and at run time.
MPILIB := $(dir $(MPICXX))../lib
$(OUT)/3dmhd_gpu: $(OUT)/obj/main_gpu.o $(OUT)/obj/gpu_hip.o $(OBJS) $(ROOT)/rast-3dmhd_amd/lib/librast3dmhd_hip.so
	g++ -O2 -fopenmp -o $@ $(OUT)/obj/main_gpu.o $(OUT)/obj/gpu_hip.o $(OBJS) \
	  $(MPILIB)/libmpicxx.so $(MPILIB)/libmpi.so -lm \
	  -L$(ROOT)/rast-3dmhd_amd/lib -lrast3dmhd_hip \
	  '-Wl,-rpath,/usr/lib/x86_64-linux-gnu:$(abspath $(MPILIB)):$$ORIGIN/../../rast-3dmhd_amd/lib'
