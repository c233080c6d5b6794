# Checker tools built from files of the reference where they lie (needs
# /root/reference; outputs only into oracle/_ref/, which git ignores).
#   exr_dump: the reference's vendored tinyexr (src/gpu/tinyexr.h) + system zlib,
#   used by tests/golden/mitsuba/make_goldens.py to decode the Mitsuba goldens.
REF ?= /root/reference
CXX ?= g++

_ref/exr_dump: exr_dump.cpp $(REF)/src/gpu/tinyexr.h
	mkdir -p _ref
	$(CXX) -O2 -std=c++17 -w -I$(REF)/src/gpu -o $@ exr_dump.cpp -lz
