# Builds the reference's own SpMV/BLAS layer (header-only sources read in place under
# the reference checkout) into oracle/_ref/, for golden vectors.  Needs the image's
# MPICH under /opt/conda; only the MPI libraries are linked from there (symlinked into
# _ref/lib so the system libstdc++ stays in use).
REF ?= /root/reference
MPI ?= /opt/conda
CXX = g++
CXXFLAGS = -std=c++14 -O3 -DNDEBUG -DXAMG_LIMIT_TYPES_USAGE -I$(REF)/src -I$(REF)/src/xamg -I$(MPI)/include

all: _ref/ref_driver

_ref/lib:
	mkdir -p _ref/lib
	for l in libmpi.so.12 libgfortran.so.4 libquadmath.so.0; do ln -sf $(MPI)/lib/$$l _ref/lib/$$l; done

_ref/ref_driver: ref_driver.cpp | _ref/lib
	$(CXX) $(CXXFLAGS) ref_driver.cpp -o $@ $(MPI)/lib/libmpi.so -Wl,-rpath,'$$ORIGIN/lib'

.PHONY: all
