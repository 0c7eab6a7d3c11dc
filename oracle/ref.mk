# Builds the reference's own host-only format converters from their sources
# in the reference tree (read in place, never copied) into oracle/_ref/.
# They are checkers for the RIPSCSR1 / RIPSADJ1 writers and the DIMACS reader:
# tests run them on small inputs and compare bytes.  Only these two files of
# the reference build standalone with g++; everything on the GPU path needs
# HIP plus a missing header (bf_hip_CSR_device_utils.hpp) or a wave32 device.
REF ?= /root/reference
SRC := $(REF)/HIP_kernel/bellman_ford/src
OUT := _ref

all: $(OUT)/gr_to_csr $(OUT)/gr_to_adjacency

$(OUT)/gr_to_csr: $(SRC)/gr_to_csr.cpp | $(OUT)
	g++ -std=c++17 -O2 $< -o $@

$(OUT)/gr_to_adjacency: $(SRC)/gr_to_adjacency.cpp | $(OUT)
	g++ -std=c++17 -O2 $< -o $@

$(OUT):
	mkdir -p $@

.PHONY: all
