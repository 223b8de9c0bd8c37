# Builds the parts of the pscbfm reference that compile from their own sources
# with plain g++, straight from where they lie under /root/reference, into
# oracle/_ref/ (git-ignored).  Only the Z-curve front study compiles without
# CUDA or LeMonADE, the header-only integer-hash RNG (rngs/Hash.h, driven by
# oracle/ref_hash_driver.cpp) and the header-only graph colouring
# (graphColoring.tpp, driven by oracle/ref_coloring_driver.cpp); the updater
# itself (.cu) is unbuildable here.
REF ?= /root/reference
OUT := $(dir $(lastword $(MAKEFILE_LIST)))_ref

all: $(OUT)/findBestZCurveFrontOrder $(OUT)/hash_ref $(OUT)/coloring_ref

$(OUT)/findBestZCurveFrontOrder: $(REF)/src/findBestZCurveFrontOrder.cpp
	mkdir -p $(OUT)
	g++ -std=c++11 -O1 -include cstdint -o $@ $<

$(OUT)/hash_ref: $(dir $(lastword $(MAKEFILE_LIST)))ref_hash_driver.cpp $(REF)/src/pscbfm/rngs/Hash.h
	mkdir -p $(OUT)
	g++ -std=c++11 -O1 -I$(REF)/src/pscbfm/rngs -o $@ $<

# -DNDEBUG as the reference's default Release flags (CMakeLists.txt:15,38-39):
# its rebalance asserts that a monomer never re-takes the colour it already
# has, which its loop can do (a no-op change) once a monomer has moved.
$(OUT)/coloring_ref: $(dir $(lastword $(MAKEFILE_LIST)))ref_coloring_driver.cpp $(REF)/src/pscbfm/graphColoring.tpp $(REF)/src/pscbfm/graphColoring.h
	mkdir -p $(OUT)
	g++ -std=c++11 -O1 -DNDEBUG -I$(REF)/src/pscbfm -o $@ $<

.PHONY: all
