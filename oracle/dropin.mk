# Builds the reference's own executables (eval/keygen.cpp, encrypt.cpp,
# cloud.cpp, decrypt.cpp under REF) against this build's idash.h and
# libraries, the way a maintainer switching the reference to this build
# would: the one change to their source is deleting the unused
# `#include <tfhe.h>`, done on the pipe into the compiler (nothing is copied).
# Outputs go to _ref/dropin/, which stays out of git. Test infrastructure
# only: tests/test_gpu_dropin.py runs them beside this build's executables
# and compares the files both write.
REF ?= /root/reference
CXX ?= g++
HOST = ../idash2019_2_amd/host
OUT = _ref/dropin
MAINS = keygen encrypt cloud decrypt

all: $(addprefix $(OUT)/,$(MAINS))

$(OUT)/%: $(REF)/eval/%.cpp $(HOST)/idash.h
	mkdir -p $(OUT)
	sed '/#include <tfhe.h>/d' $< | $(CXX) -O2 -std=c++17 -fopenmp -I$(HOST) -I../include -x c++ - \
	  -o $@ -L../idash2019_2_amd -lidash_host -lidash_hip -Wl,-rpath,'$$ORIGIN/../../../idash2019_2_amd'

clean:
	rm -rf $(OUT)

.PHONY: all clean
