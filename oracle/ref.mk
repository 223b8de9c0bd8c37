# Builds the parts of the pscbfm reference that compile from their own sources
# with plain g++, straight from where they lie under /root/reference, into
# oracle/_ref/ (git-ignored).  Only the Z-curve front study compiles without
# CUDA or LeMonADE; the updater itself (.cu) is unbuildable here.
REF ?= /root/reference
OUT := $(dir $(lastword $(MAKEFILE_LIST)))_ref

all: $(OUT)/findBestZCurveFrontOrder

$(OUT)/findBestZCurveFrontOrder: $(REF)/src/findBestZCurveFrontOrder.cpp
	mkdir -p $(OUT)
	g++ -std=c++11 -O1 -include cstdint -o $@ $<

.PHONY: all
