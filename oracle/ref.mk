# Builds the reference's own CD / CD-RMP2 / CD-UMP2 / CD-RMP3 / CD-UMP3 code from its sources
# (read where they lie under the reference checkout) with its vendored Eigen,
# plus the driver in this directory.  Output goes to oracle/_ref/ only.
# IntegralEngine (libint2) is not built; the one symbol of it that the MP2
# translation units reference belongs to constructors the driver never calls.
REF  ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT  := $(HERE)_ref
SRCS := $(REF)/src/integrals/cholesky_eri.cc $(REF)/src/mp2/cholesky_rmp2.cc \
        $(REF)/src/mp2/cholesky_ump2.cc $(REF)/src/mp3/cholesky_rmp3.cc $(REF)/src/mp3/cholesky_ump3.cc \
        $(REF)/src/core/molecule.cc $(REF)/src/core/basis.cc
CXXFLAGS := -std=c++17 -O2 -fopenmp -I$(REF)/include -I$(REF)/include/eigen-3.4.0 -w

$(OUT)/ref_driver: $(HERE)ref_driver.cc $(SRCS)
	mkdir -p $(OUT)
	g++ $(CXXFLAGS) -o $@ $(HERE)ref_driver.cc $(SRCS) -Wl,--unresolved-symbols=ignore-all

.PHONY: clean
clean:
	rm -rf $(OUT)
