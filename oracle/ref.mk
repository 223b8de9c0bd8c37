# Builds the parts of the pscbfm reference that compile from their own sources
# with plain g++, straight from where they lie under /root/reference, into
# oracle/_ref/ (git-ignored).  Only the Z-curve front study compiles without
# CUDA or LeMonADE, and the header-only integer-hash RNG (rngs/Hash.h, driven
# by oracle/ref_hash_driver.cpp); the updater itself (.cu) is unbuildable here.
REF ?= /root/reference
OUT := $(dir $(lastword $(MAKEFILE_LIST)))_ref

all: $(OUT)/findBestZCurveFrontOrder $(OUT)/hash_ref

$(OUT)/findBestZCurveFrontOrder: $(REF)/src/findBestZCurveFrontOrder.cpp
	mkdir -p $(OUT)
	g++ -std=c++11 -O1 -include cstdint -o $@ $<

$(OUT)/hash_ref: $(dir $(lastword $(MAKEFILE_LIST)))ref_hash_driver.cpp $(REF)/src/pscbfm/rngs/Hash.h
	mkdir -p $(OUT)
	g++ -std=c++11 -O1 -I$(REF)/src/pscbfm/rngs -o $@ $<

.PHONY: all
