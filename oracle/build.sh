#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/_build/.
# Plain IEEE single precision: no FMA contraction, no fast-math.
set -e
cd "$(dirname "$0")"
mkdir -p _build
gcc -std=c11 -O2 -fPIC -shared -fopenmp -ffp-contract=off -fno-fast-math \
    -Wall -Wextra -Wno-unused-parameter \
    -o _build/libslosh_oracle.so slosh_oracle.c -lm
