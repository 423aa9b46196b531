# Builds the parts of the reference that compile from their own sources with
# g++ alone (no CUDA): the .prm parser Src/FileParams.cpp, driven by
# ref_fileparams_main.cpp.  Outputs go to oracle/_ref/ (git-ignored).
# Every other reference source includes cuda_runtime.h or uses CUDA textures
# and cannot be built here.
REF ?= /root/reference
OUT := _ref

all: $(OUT)/ref_fileparams

$(OUT)/ref_fileparams: ref_fileparams_main.cpp $(REF)/Src/FileParams.cpp $(REF)/Include/FileParams.h
	mkdir -p $(OUT)
	g++ -O1 -w -I$(REF)/Include -o $@ ref_fileparams_main.cpp $(REF)/Src/FileParams.cpp
