#!/bin/bash
# Builds oracle/_ref/rang_lists: the reference's own angular-coefficient
# producer (GRASP2018 libmod + lib9290 + librang90, and rci90's TALK/CXK
# callbacks) compiled from the sources where they lie under
# /root/reference (read-only), linked with oracle/rang_driver.f90.  Test
# infrastructure only: it writes the per-pair lists rci90's SETHAM consumes
# (tests/golden/make_rang_golden.py turns them into fixtures).  Nothing under
# /root/reference is copied.  Needs amdflang (ROCm); exits 0 without
# building when the reference tree or the compiler is absent.
#
# The libraries' module dependencies are resolved by compiling in passes:
# every source that fails for want of a module compiled later is retried on
# the next pass, until a pass makes no progress.  The link goes through a
# static archive, so only the objects the driver reaches are pulled in (the
# LAPACK-dependent ones, e.g. INIEST2 and SETMC, are not).
set -uo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${GRASP_REFERENCE:-/root/reference}/src"
FC="${FC:-/opt/rocm/bin/amdflang}"
OUT="$HERE/_ref"
if [ ! -d "$REF" ] || [ ! -x "$FC" ]; then
  echo "build_rang: reference tree or amdflang absent; skipping" >&2
  exit 0
fi
OBJ="$OUT/rang_obj"
rm -rf "$OBJ"
mkdir -p "$OBJ"
cd "$OBJ"
pending=()
for d in lib/libmod lib/lib9290 lib/librang90; do
  for f in "$REF/$d"/*.f90; do pending+=("$f"); done
done
for f in talk_I cxk_I talk cxk breid_I breid; do pending+=("$REF/appl/rci90/$f.f90"); done
export FC
compile_one() {
  f="$1"
  o="$(basename "$(dirname "$f")")_$(basename "$f" .f90).o"
  "$FC" -O1 -fPIC -c "$f" -o "$o" > /dev/null 2>&1 || rm -f "$o"
}
export -f compile_one
pass=0
while [ ${#pending[@]} -gt 0 ]; do
  pass=$((pass + 1))
  printf '%s\n' "${pending[@]}" | xargs -P "${BUILD_JOBS:-8}" -I{} bash -c 'compile_one "$@"' _ {}
  still=()
  for f in "${pending[@]}"; do
    o="$(basename "$(dirname "$f")")_$(basename "$f" .f90).o"
    [ -f "$o" ] || still+=("$f")
  done
  if [ ${#still[@]} -eq ${#pending[@]} ]; then
    echo "build_rang: ${#still[@]} sources do not compile:" >&2
    printf '  %s\n' "${still[@]}" >&2
    exit 1
  fi
  pending=("${still[@]}")
done
ar rcs librang_ref.a ./*.o
"$FC" -O1 -c "$HERE/rang_driver.f90" -o rang_driver.o
"$FC" -o "$OUT/rang_lists" rang_driver.o librang_ref.a
cd "$HERE"
rm -rf "$OBJ"
echo "build_rang: $OUT/rang_lists ($pass passes)"
