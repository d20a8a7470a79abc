# Builds oracle/_ref/libsimall_ref.so from the reference's own source files
# where they lie under /root/reference, plus oracle/ref_shim.cpp (flat C entry
# points).  Only in a container that has the reference; outputs go to _ref/.
# TEST INFRASTRUCTURE: the checker's checker, never part of the product.
REF ?= /root/reference
CXX ?= g++
OUT := _ref
SRCS := $(REF)/src/meshing/CartesianMesher.cpp $(REF)/src/meshing/Connectivity.cpp \
        $(REF)/src/meshing/MeshStorage.cpp $(REF)/src/solver/SimpleAlgorithm.cpp \
        $(REF)/src/solver/Gradient.cpp $(REF)/src/solver/PisoAlgorithm.cpp \
        $(REF)/src/adjoint/DiscreteAdjoint.cpp \
        $(REF)/src/solver/AMGPreconditioner.cpp $(REF)/src/io/Checkpoint.cpp
# -ffp-contract=off: x86-64 code without fused multiply-add, as the
# reference's default build.
FLAGS := -std=c++20 -O2 -ffp-contract=off -fPIC -I$(REF)/src -I$(REF) -w

all: $(OUT)/libsimall_ref.so

$(OUT)/libsimall_ref.so: ref_shim.cpp $(SRCS)
	mkdir -p $(OUT)
	$(CXX) $(FLAGS) -shared -o $@.tmp ref_shim.cpp $(SRCS) && mv $@.tmp $@

clean:
	rm -rf $(OUT)
.PHONY: all clean
