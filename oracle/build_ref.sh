#!/bin/sh
# Compile the reference's FFT-free hot-path subroutines (tridv, complen,
# compsgs1, sufsim/lower, sufsim_upper/upper, compmn, and the modules and
# allocation they need) from /root/reference, promoted to FP64, together with
# the test driver oracle/ref_harness.f90.  Output only into oracle/_ref/.
# The FFT-dependent files (filter_fftw.f90 and its callers) need FFTW 2.x,
# which this image does not have, so they are not built.
set -e
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-flang}
command -v "$FC" >/dev/null 2>&1 || FC=/opt/rocm/lib/llvm/bin/flang
[ -d "$REF" ] || { echo "no reference at $REF" >&2; exit 2; }
mkdir -p "$OUT/obj"
FFLAGS="-O2 -fdefault-real-8 -fdefault-double-8 -J $OUT/obj"
for f in parameters.f90 modules.f90 compbuf.f90 buffers.f90 allocate.f90 default.f90 \
         abort.F90 tridv.f90 complen.f90 compsgs1.f90 lowerbc.f90 upperbc.f90 compmn.F90; do
  $FC $FFLAGS -c "$REF/$f" -o "$OUT/obj/${f%.*}.o"
done
$FC $FFLAGS -c "$HERE/ref_harness.f90" -o "$OUT/obj/ref_harness.o"
$FC -o "$OUT/ref_harness" "$OUT"/obj/*.o
echo "built $OUT/ref_harness"
