# Builds the parts of the reference that compile from their own sources with
# g++ alone (no CUDA): the .prm parser Src/FileParams.cpp, driven by
# ref_fileparams_main.cpp.  Outputs go to oracle/_ref/ (git-ignored).
# Every other reference source includes cuda_runtime.h or uses CUDA textures
# and cannot be built here.
#
# ref_main: the reference's own CLI, Src/main.cpp, compiled unchanged against
# this package's drop-in headers (include/listctrecons) and linked to its
# libraries in place of the reference's CUDA objects (makefile:4-18), which
# is the drop-in a maintainer makes.  Needs the package built first.
REF ?= /root/reference
OUT := _ref
PKG := ../listctrecons_amd

all: $(OUT)/ref_fileparams

$(OUT)/ref_main: $(REF)/Src/main.cpp $(PKG)/libListCTRecons.so $(PKG)/liblct.so
	mkdir -p $(OUT)
	g++ -O1 -std=c++17 -I../include -I../include/listctrecons -o $@ $(REF)/Src/main.cpp \
	    -L$(PKG) -lListCTRecons -llct '-Wl,-rpath,$$ORIGIN/../../listctrecons_amd'

$(OUT)/ref_fileparams: ref_fileparams_main.cpp $(REF)/Src/FileParams.cpp $(REF)/Include/FileParams.h
	mkdir -p $(OUT)
	g++ -O1 -w -I$(REF)/Include -o $@ ref_fileparams_main.cpp $(REF)/Src/FileParams.cpp
