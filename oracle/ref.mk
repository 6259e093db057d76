# Builds the reference Velox CPU library directly from its own sources under
# /root/reference (no reference build system is run) into oracle/_ref/.
# Two flavours: "strict" (scalar math, no contraction: the parity oracle) and
# "relaxed" (the reference's default -O3 Release flags: the timed CPU baseline).
REF      ?= /root/reference
OUT      ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CXX      ?= g++
SRCS     := world solver queries debug character ragdoll vehicle serialization \
            serialization_v2 rollback benchmark velox_c stable_api manifold sleep softbody
COMMON   := -std=c++17 -O3 -DNDEBUG -fPIC -I$(REF)/include -I$(REF)/src \
            -DVELOX_HAS_CUDA=0 -DVELOX_HAS_VULKAN=0 -fno-math-errno
STRICT   := $(COMMON) -DVELOX_STRICT_FLOATING_POINT=1 -fno-fast-math -ffp-contract=off
RELAXED  := $(COMMON) -DVELOX_STRICT_FLOATING_POINT=0 -ffp-contract=fast

STRICT_OBJS  := $(addprefix $(OUT)/strict/,$(addsuffix .o,$(SRCS)))
RELAXED_OBJS := $(addprefix $(OUT)/relaxed/,$(addsuffix .o,$(SRCS)))

all: $(OUT)/libvelox_strict.a $(OUT)/libvelox_relaxed.a

$(OUT)/strict/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(STRICT) -c $< -o $@

$(OUT)/relaxed/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(RELAXED) -c $< -o $@

$(OUT)/libvelox_strict.a: $(STRICT_OBJS)
	ar rcs $@ $^

$(OUT)/libvelox_relaxed.a: $(RELAXED_OBJS)
	ar rcs $@ $^

.PHONY: all

# Harness executables (oracle/ref_harness.cpp): golden-vector writer against the
# strict library, CPU-baseline timer against the relaxed one.
HARNESS := $(dir $(lastword $(MAKEFILE_LIST)))ref_harness.cpp
HDEPS   := $(HARNESS) $(dir $(lastword $(MAKEFILE_LIST)))../velox_amd/csrc/scene_gen.hpp

$(OUT)/ref_fixtures: $(HDEPS) $(OUT)/libvelox_strict.a
	$(CXX) $(STRICT) $(HARNESS) $(OUT)/libvelox_strict.a -pthread -o $@

$(OUT)/ref_bench: $(HDEPS) $(OUT)/libvelox_relaxed.a
	$(CXX) $(RELAXED) $(HARNESS) $(OUT)/libvelox_relaxed.a -pthread -o $@

# The HipBackend adapter installed in the strict reference World, linked with
# this repository's libvxhip.so (built first by velox_amd/build.py).
REPO    := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))..)
$(OUT)/ref_hip: $(HDEPS) $(dir $(lastword $(MAKEFILE_LIST)))ref_hip.cpp $(OUT)/libvelox_strict.a $(REPO)/include/vxhip.h
	$(CXX) $(STRICT) -I$(REPO)/include $(dir $(lastword $(MAKEFILE_LIST)))ref_hip.cpp $(OUT)/libvelox_strict.a \
	    -L$(REPO)/velox_amd/_build -lvxhip -Wl,-rpath,'$$ORIGIN/../../velox_amd/_build' -pthread -o $@

harness: $(OUT)/ref_fixtures $(OUT)/ref_bench
hip: $(OUT)/ref_hip
.PHONY: harness hip
