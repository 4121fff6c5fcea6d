# Host C/C++ of the set-up path built with AddressSanitizer and
# UndefinedBehaviorSanitizer (gcc, host code only; no device code is built
# here), for tests/test_host_sanitized.py:
#   libmpk_host_san.so   mojo-petsc4py_amd/csrc/mpk_factor.cpp (symbolic and
#                        numeric factorisations, level sets, sweep planner)
#   liboracle_san.so     oracle/oracle.c (the checker)
# Arithmetic flags as in the regular builds (-ffp-contract=off), so results
# stay bitwise comparable.
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
ROOT := $(abspath $(HERE)..)
OUT := $(HERE)_san_build
SAN := -fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer -g -O1

all: $(OUT)/libmpk_host_san.so $(OUT)/liboracle_san.so

$(OUT)/libmpk_host_san.so: $(ROOT)/mojo-petsc4py_amd/csrc/mpk_factor.cpp $(ROOT)/include/mpk.h
	mkdir -p $(OUT)
	g++ $(SAN) -std=c++17 -fPIC -shared -ffp-contract=off -o $@ $<

$(OUT)/liboracle_san.so: $(ROOT)/oracle/oracle.c
	mkdir -p $(OUT)
	gcc $(SAN) -std=c11 -fPIC -shared -ffp-contract=off -fno-fast-math -o $@ $< -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
