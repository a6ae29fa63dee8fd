# Builds the parts of the reference that compile from their own sources,
# unchanged, where they lie, into oracle/_ref/ (test infrastructure only:
# the checker of the oracle's restatement, never shipped or measured).
#   make -f ref.mk      -> oracle/_ref/librng.so   (src/rng.c, src/rng.h)
# Everything else of the reference path is CUDA (nvcc and a CUDA device)
# or needs MPI/CGNS headers the image lacks, so it is not built here.
REF ?= /root/reference/src
CC ?= gcc

all: _ref/librng.so

_ref/librng.so: $(REF)/rng.c $(REF)/rng.h
	mkdir -p _ref
	$(CC) -O2 -fPIC -shared -I$(REF) -o $@ $(REF)/rng.c -lm

clean:
	rm -rf _ref

.PHONY: all clean
