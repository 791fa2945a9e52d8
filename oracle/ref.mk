# Builds the reference's own CPU code for this path, unmodified, from the
# sources where they lie, into oracle/_ref/ (git-ignored). Used only as a
# checker: to generate golden vectors and to pin the oracle restatement.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# The flags are the ones the reference itself uses:
#   reversible/free_dp.py:193-196   g++ -O3 -std=c++17 -fopenmp <src> -o relative5d
#   reversible/world_dp.py:154-175  g++ -O3 -std=c++17 -fopenmp <src> -o world8d
#   local_game/native.py:21-34      g++ -O3 -std=c++17 -shared -fPIC -fopenmp <src>
REF ?= /root/reference/src/pursuit_evasion
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref

all: $(OUT)/relative5d $(OUT)/libnative_geometry.so $(OUT)/world8d

$(OUT)/relative5d: $(REF)/reversible/native_relative.cpp
	mkdir -p $(OUT)
	g++ -O3 -std=c++17 -fopenmp $< -o $@

$(OUT)/world8d: $(REF)/reversible/native_world.cpp
	mkdir -p $(OUT)
	g++ -O3 -std=c++17 -fopenmp $< -o $@

$(OUT)/libnative_geometry.so: $(REF)/local_game/native_geometry.cpp
	mkdir -p $(OUT)
	g++ -O3 -std=c++17 -shared -fPIC -fopenmp $< -o $@

.PHONY: all
