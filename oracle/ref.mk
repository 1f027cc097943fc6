# Builds the reference FargoCPT sources, where they lie under /root/reference/src,
# into object files and static libraries under oracle/_ref/ (never into the
# reference tree, never into the repository history).
#
#   make -f oracle/ref.mk -j8            # from the repository root
#
# Build flags differ from the reference's own makefile.defs on purpose:
#   -O2 -fno-fast-math -ffp-contract=off   (its default is -Ofast -march=native -flto,
#                                          which reassociates the sequential ring
#                                          sums and contracts FMAs by host ISA)
#   -DDISABLE_FFTW -DDISABLE_GSL           (both libraries are absent from this image;
#                                          the reference guards every use with these)
# MPI comes from the MPICH that ships in this image (/opt/conda).
REF      ?= /root/reference/src
OUT      ?= oracle/_ref
MPI_HOME ?= /opt/conda
CXX      := g++
CC       := gcc

FLAGS_COMMON := -O2 -fno-fast-math -ffp-contract=off -g0 -fopenmp
CXXFLAGS := $(FLAGS_COMMON) -std=c++17 -DNDEBUG -DOMPI_SKIP_MPICXX -DDISABLE_FFTW -DDISABLE_GSL \
            -I$(REF) -I$(MPI_HOME)/include -w
CFLAGS_RB := -O2 -fno-fast-math -ffp-contract=off -std=c99 -D_GNU_SOURCE -w
CXXFLAGS_YAML := -O2 -fPIC -std=gnu++11 -I$(REF) -w
CXXFLAGS_UNITS := -O2 -fPIC -std=gnu++17 -I$(REF) -w

FARGO_SRCS := $(notdir $(wildcard $(REF)/*.cpp)) \
              $(addprefix particles/,$(notdir $(wildcard $(REF)/particles/*.cpp))) \
              $(addprefix random/,$(notdir $(wildcard $(REF)/random/*.cpp))) \
              $(addprefix nbody/,$(notdir $(wildcard $(REF)/nbody/*.cpp))) \
              $(addprefix viscosity/,$(notdir $(wildcard $(REF)/viscosity/*.cpp))) \
              $(addprefix boundary_conditions/,$(notdir $(wildcard $(REF)/boundary_conditions/*.cpp)))
FARGO_OBJS := $(FARGO_SRCS:%.cpp=$(OUT)/fargo/%.o)
RB_SRCS    := rebound.c tree.c particle.c gravity.c integrator.c integrator_whfast.c integrator_saba.c \
              integrator_ias15.c integrator_sei.c integrator_bs.c integrator_leapfrog.c integrator_mercurius.c \
              integrator_eos.c integrator_tes.c boundary.c input.c binarydiff.c output.c collision.c \
              communication_mpi.c display.c tools.c rotations.c derivatives.c simulationarchive.c glad.c \
              integrator_janus.c transformations.c
RB_OBJS    := $(RB_SRCS:%.c=$(OUT)/rebound/%.o)
YAML_OBJS  := $(patsubst $(REF)/yaml-cpp/%.cpp,$(OUT)/yaml-cpp/%.o,$(wildcard $(REF)/yaml-cpp/*.cpp))
UNITS_OBJS := $(patsubst $(REF)/units/%.cpp,$(OUT)/units/%.o,$(wildcard $(REF)/units/*.cpp))

all: $(OUT)/libfargo_ref.a $(OUT)/librebound.a $(OUT)/libyamlcpp.a $(OUT)/libunits.a $(OUT)/fargocpt_ref

$(OUT)/fargo/%.o: $(REF)/%.cpp
	@mkdir -p $(@D)
	$(CXX) -c $(CXXFLAGS) -o $@ $<

# stress.cpp calls PersonalExit() inside its DISABLE_FFTW branch without including the
# reference header that declares it; force-include that same reference header.
$(OUT)/fargo/stress.o: $(REF)/stress.cpp
	@mkdir -p $(@D)
	$(CXX) -c $(CXXFLAGS) -include $(REF)/LowTasks.h -o $@ $<

$(OUT)/rebound/%.o: $(REF)/rebound/%.c
	@mkdir -p $(@D)
	$(CC) -c $(CFLAGS_RB) -o $@ $<

$(OUT)/yaml-cpp/%.o: $(REF)/yaml-cpp/%.cpp
	@mkdir -p $(@D)
	$(CXX) -c $(CXXFLAGS_YAML) -o $@ $<

$(OUT)/units/%.o: $(REF)/units/%.cpp
	@mkdir -p $(@D)
	$(CXX) -c $(CXXFLAGS_UNITS) -o $@ $<

# everything except main.o goes into the library the fixture driver links
$(OUT)/libfargo_ref.a: $(filter-out $(OUT)/fargo/main.o,$(FARGO_OBJS))
	ar rcs $@ $^
$(OUT)/librebound.a: $(RB_OBJS)
	ar rcs $@ $^
$(OUT)/libyamlcpp.a: $(YAML_OBJS)
	ar rcs $@ $^
$(OUT)/libunits.a: $(UNITS_OBJS)
	ar rcs $@ $^

# libmpi is linked through a directory that holds only a link to it: putting
# $(MPI_HOME)/lib itself on the search path would pull in that directory's older
# libstdc++ in place of the system one.
MPILINK := $(abspath $(OUT))/mpilib
LIBS := $(OUT)/libfargo_ref.a $(OUT)/librebound.a $(OUT)/libyamlcpp.a $(OUT)/libunits.a \
        -L$(MPILINK) -Wl,-rpath,$(MPILINK) -lmpi -lm -lrt -fopenmp

$(MPILINK)/libmpi.so:
	@mkdir -p $(MPILINK)
	ln -sf $(MPI_HOME)/lib/libmpi.so.12 $(MPILINK)/libmpi.so.12
	ln -sf $(MPI_HOME)/lib/libmpi.so.12 $(MPILINK)/libmpi.so
	ln -sf $(MPI_HOME)/lib/libgfortran.so.4 $(MPILINK)/libgfortran.so.4
	ln -sf $(MPI_HOME)/lib/libquadmath.so.0 $(MPILINK)/libquadmath.so.0

# the unmodified reference executable (used for the speed comparison with the C restatement)
$(OUT)/fargocpt_ref: $(MPILINK)/libmpi.so $(OUT)/fargo/main.o $(OUT)/libfargo_ref.a $(OUT)/librebound.a $(OUT)/libyamlcpp.a $(OUT)/libunits.a
	$(CXX) -o $@ $(OUT)/fargo/main.o $(LIBS)

.PHONY: all
