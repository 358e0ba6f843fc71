#!/bin/bash
# ORACLE -- test infrastructure only.
#
# Builds the reference's CPU VM from the reference sources where they lie
# into oracle/_ref/libvirgref.so, with oracle/ref_harness.c as the driver.
# Nothing of the reference is copied: the sources are compiled in place and
# only objects and the library land in oracle/_ref/ (git-ignored).
#
# Compiled: util.c, tablet/, table/, db/, reader/, vm/ (virginia.c twice,
# single- and multi-core, as the reference's own Makefile does), release.c.
# Left out, and why:
#   init.c, close.c     need libcudart (absent); ref_harness.c does their
#                       non-CUDA part
#   vm/execute.c        calls virg_vm_gpu, whose source (vm_gpu.cu) needs nvcc
#                       (absent); ref_harness.c walks the same serial ops and
#                       calls the reference's virg_vm_cpu
#   query.c, sql/*      the SQL grammar needs bison and flex (absent)
#   opcodelist.c        generated at build time; only printing uses it
# virginian.h includes CUDA's runtime headers for declarations only; the
# image carries them inside the triton package.
#
#   oracle/build_ref.sh [reference-src-dir]
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${1:-/root/reference/src}
OUT=$HERE/_ref
if [ ! -f "$REF/virginian.h" ]; then
	echo "build_ref.sh: no reference sources at $REF; nothing built" >&2
	exit 0
fi
CUDA_INC=${CUDA_INC:-$(python3 -c 'import os, triton; print(os.path.join(os.path.dirname(triton.__file__), "backends", "nvidia", "include"))' 2>/dev/null || true)}
if [ -z "$CUDA_INC" ] || [ ! -f "$CUDA_INC/cuda_runtime_api.h" ]; then
	echo "build_ref.sh: CUDA runtime headers not found; the reference VM is unbuildable here" >&2
	exit 0
fi
mkdir -p "$OUT/obj"
CFLAGS="-std=gnu99 -O2 -fPIC -fcommon -w -fvisibility=hidden -ffunction-sections -fdata-sections -pthread"
INC="-I$REF -I$REF/sql -I$CUDA_INC"
SRCS="util.c release.c"
for d in tablet table db reader; do
	for f in "$REF"/$d/*.c; do SRCS="$SRCS $d/$(basename "$f")"; done
done
SRCS="$SRCS vm/addop.c vm/allocresult.c vm/cleanup.c vm/freeresults.c vm/init.c vm/vm_cpu.c vm/virginia.c"
OBJS=""
for s in $SRCS; do
	o="$OUT/obj/$(echo "$s" | tr '/' '_' | sed 's/\.c$/.o/')"
	gcc $CFLAGS $INC -c "$REF/$s" -o "$o"
	OBJS="$OBJS $o"
done
gcc $CFLAGS $INC -D __SINGLE -c "$REF/vm/virginia.c" -o "$OUT/obj/vm_virginia_single.o"
gcc $CFLAGS $INC -c "$HERE/ref_harness.c" -o "$OUT/obj/ref_harness.o"
gcc -shared -Wl,--gc-sections -Wl,--no-undefined -pthread -o "$OUT/libvirgref.so" \
	$OBJS "$OUT/obj/vm_virginia_single.o" "$OUT/obj/ref_harness.o" -lm
echo "$OUT/libvirgref.so"
