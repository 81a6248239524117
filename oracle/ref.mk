# Builds the reference's INI reader (three standalone C++ files of
# /root/reference/src/config) plus oracle/ref_ini_shim.cpp into
# oracle/_ref/libref_ini.so.  Reference sources are compiled where they lie;
# nothing from the reference is copied.  The reference's advection/conv code
# needs a SYCL compiler and the mdspan submodule, which this image lacks, so
# it is not built.
REF     ?= /root/reference
CFG     := $(REF)/src/config
HERE    := $(abspath $(dir $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)/_ref/libref_ini.so

all: $(OUT)

$(OUT): $(HERE)/ref_ini_shim.cpp $(CFG)/ConfigMap.cpp $(CFG)/inih/INIReader.cpp $(CFG)/inih/ini.cpp
	@mkdir -p $(HERE)/_ref
	g++ -O2 -fPIC -shared -std=c++17 -w -I$(CFG) -I$(CFG)/inih -o $@ $^

clean:
	rm -rf $(HERE)/_ref

.PHONY: all clean
