#!/bin/sh
# Build the CPU oracle (test infrastructure only) with the reference's release
# flags (src/fvm/etc/buildsystem/MyTools/g++.py:31) and no FMA contraction.
#   sh oracle/build.sh            -> oracle/liboracle.so
#   sh oracle/build.sh sanitize   -> oracle/_sanitize/sanitize_main (AddressSanitizer +
#                                    UndefinedBehaviorSanitizer build of the oracle sources)
set -e
here=$(cd "$(dirname "$0")" && pwd)
if [ "$1" = "sanitize" ]; then
    mkdir -p "$here/_sanitize"
    # built under a private name and renamed into place, so that a concurrent
    # build never replaces the file while another process starts it
    tmp="$here/_sanitize/sanitize_main.$$"
    gcc -O1 -g -fno-omit-frame-pointer -ffp-contract=off -fsanitize=address,undefined \
        -fno-sanitize-recover=all -Wall -o "$tmp" \
        "$here/sanitize_main.c" "$here/amg_oracle.c" "$here/fv_oracle.c" -lm
    mv -f "$tmp" "$here/_sanitize/sanitize_main"
    exit 0
fi
gcc -O3 -finline-limit=500 -ffp-contract=off -fPIC -shared -Wall -o "$here/liboracle.so" \
    "$here/amg_oracle.c" "$here/fv_oracle.c" -lm
