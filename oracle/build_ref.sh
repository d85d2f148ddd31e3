#!/bin/bash
# Build the reference's own kernels, from their sources where they lie, into
# oracle/_ref/ (git-ignored).  Test infrastructure only: the outputs are used
# by tests/ to pin the oracle, by tests/golden/gen_golden.py to write the
# fixtures, and by bench.py's cpu_baseline leg.  Nothing here is shipped.
#
#   _ref/libref_ieee.so  every kernel function, -O2 -fno-fast-math
#                        -ffp-contract=off: the IEEE source-order build that the
#                        oracle must match bit for bit.
#   _ref/libref_fast.so  the same kernels with the reference's own flags
#                        (makefile.in:28-29: -O3 -ffast-math -march=...;
#                        whispering/gcc/makefile:24: -g -O0).  Objects are
#                        compiled with -ffast-math but linked without it, so
#                        loading the library does not switch the process to
#                        flush-to-zero.
#   _ref/bin/<test>      the reference's gcc-target programs (main included),
#                        reference flags, for its printed means.
#   _ref/libref_omp.so   the C kernels with the reference flags plus -fopenmp
#                        (their `#pragma omp parallel for` active): the
#                        multi-core CPU baseline of bench.py.  The Fortran
#                        kernels are left out (a second OpenMP runtime).
#
# Each C test file is compiled with -Dmain=kgref_main_<test> so that its kernel
# function can be called directly; nothing is copied or stubbed.  The Fortran
# kernels (jacobi, matmul, sincos) are built with amdflang, gfortran being
# absent.  Every test is built as the reference builds it: real = double
# (makefile.in:12, which overrides the `real ?= float` of the vecadd, matvec
# and sincos makefiles) and -fdefault-real-8 for Fortran.
#
# -march defaults to x86-64-v4 (AVX-512) rather than native so that the
# prebuilt files run on any current x86-64 host; set REF_MARCH=native to match
# makefile.in exactly on the machine that runs them.
set -euo pipefail

REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
MARCH=${REF_MARCH:-x86-64-v4}

if [ ! -d "$REF" ]; then
    echo "build_ref.sh: $REF not present; skipping the reference build" >&2
    exit 0
fi

CTESTS="laplacian divergence gradient wave13pt uxx1 tricubic tricubic2 lapgsrb gaussblur whispering gameoflife vecadd matvec"
FTESTS="jacobi matmul sincos"

rm -rf "$OUT"
mkdir -p "$OUT/ieee" "$OUT/fast" "$OUT/omp" "$OUT/bin"

COMMON="-std=c99 -D_GNU_SOURCE -Dreal=double -fPIC -w"
IEEE="-O2 -fno-fast-math -ffp-contract=off"
FAST="-O3 -ffast-math -march=$MARCH"
FAST_O0="-g -O0"

gcc $COMMON $IEEE -c "$REF/laplacian/timing.c" -o "$OUT/ieee/timing.o"
gcc $COMMON $FAST -c "$REF/laplacian/timing.c" -o "$OUT/fast/timing.o"

for t in $CTESTS; do
    src="$REF/$t/$t.c"
    gcc $COMMON $IEEE -I"$REF/$t" -Dmain=kgref_main_$t -c "$src" -o "$OUT/ieee/$t.o"
    if [ "$t" = whispering ]; then fl="$FAST_O0"; else fl="$FAST"; fi
    gcc $COMMON $fl -I"$REF/$t" -Dmain=kgref_main_$t -c "$src" -o "$OUT/fast/$t.o"
    if [ "$t" != whispering ]; then
        gcc $COMMON $FAST -fopenmp -I"$REF/$t" -Dmain=kgref_main_$t -c "$src" -o "$OUT/omp/$t.o"
    fi
    # The reference gcc-target program itself (timing.c from the same test dir).
    gcc $COMMON $fl -I"$REF/$t" "$src" "$REF/$t/timing.c" -o "$OUT/bin/$t" -lrt
done

for t in $FTESTS; do
    amdflang -cpp -fdefault-real-8 -fPIC $IEEE -c "$REF/$t/$t.F90" -o "$OUT/ieee/${t}_f.o"
    amdflang -cpp -fdefault-real-8 -fPIC $FAST -c "$REF/$t/$t.F90" -o "$OUT/fast/${t}_f.o"
    gcc $COMMON $IEEE -I"$REF/$t" -Dmain=kgref_main_$t -c "$REF/$t/main.c" -o "$OUT/ieee/${t}_main.o"
    gcc $COMMON $FAST -I"$REF/$t" -c "$REF/$t/main.c" -o "$OUT/fast/${t}_prog.o"
    gcc "$OUT/fast/${t}_prog.o" "$OUT/fast/${t}_f.o" "$OUT/fast/timing.o" -o "$OUT/bin/$t" -lrt -lm
done

gcc -shared -o "$OUT/libref_ieee.so" "$OUT"/ieee/*.o -lrt -lm
gcc -shared -o "$OUT/libref_fast.so" $(ls "$OUT"/fast/*.o | grep -v _prog.o) -lrt -lm
gcc -shared -fopenmp -o "$OUT/libref_omp.so" "$OUT"/omp/*.o "$OUT/fast/timing.o" -lrt -lm
rm -rf "$OUT/ieee" "$OUT/fast" "$OUT/omp"
echo "build_ref.sh: built $OUT (march=$MARCH)"
