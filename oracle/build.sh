#!/bin/sh
# Build the CPU oracle (test infrastructure only) with the reference's release
# flags (src/fvm/etc/buildsystem/MyTools/g++.py:31) and no FMA contraction.
set -e
here=$(cd "$(dirname "$0")" && pwd)
gcc -O3 -finline-limit=500 -ffp-contract=off -fPIC -shared -Wall -o "$here/liboracle.so" \
    "$here/amg_oracle.c" "$here/fv_oracle.c" -lm
