# Builds the reference PIAFS solver from its own sources, in place, into
# oracle/_ref/ (git-ignored).  Serial build (-Dserial) plus an OpenMP variant.
# The reference tree is read-only; objects go under oracle/_ref/obj*.
# Usage: make -f oracle/ref.mk -j8   (from the repository root)
REF      ?= /root/reference
OUT      := oracle/_ref
CSRC     := $(shell find $(REF)/src -name '*.c')
CXXSRC   := $(shell find $(REF)/src -name '*.cpp')
COBJ     := $(patsubst $(REF)/src/%.c,$(OUT)/obj/%.o,$(CSRC))
CXXOBJ   := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/%.o,$(CXXSRC))
COBJ_OMP   := $(patsubst $(REF)/src/%.c,$(OUT)/obj_omp/%.o,$(CSRC))
CXXOBJ_OMP := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj_omp/%.o,$(CXXSRC))
CFLAGS   := -std=c99 -O2 -Dserial -I$(REF)/include -Wno-alloc-size-larger-than -w
CXXFLAGS := -std=c++11 -O2 -Dserial -I$(REF)/include -w

all: $(OUT)/PIAFS $(OUT)/PIAFS_omp $(OUT)/CompareSolutions

$(OUT)/PIAFS: $(COBJ) $(CXXOBJ)
	g++ -o $@ $^ -lm
$(OUT)/PIAFS_omp: $(COBJ_OMP) $(CXXOBJ_OMP)
	g++ -fopenmp -o $@ $^ -lm
$(OUT)/CompareSolutions: $(REF)/Extras/CompareSolutions.c
	@mkdir -p $(dir $@)
	gcc -std=c99 -O2 -w -o $@ $< -lm

$(OUT)/obj/%.o: $(REF)/src/%.c
	@mkdir -p $(dir $@)
	gcc $(CFLAGS) -c $< -o $@
$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	g++ $(CXXFLAGS) -c $< -o $@
$(OUT)/obj_omp/%.o: $(REF)/src/%.c
	@mkdir -p $(dir $@)
	gcc $(CFLAGS) -fopenmp -Dwith_omp -c $< -o $@
$(OUT)/obj_omp/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	g++ $(CXXFLAGS) -fopenmp -Dwith_omp -c $< -o $@

.PHONY: all
