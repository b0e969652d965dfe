#!/bin/bash
# Builds oracle/_ref/libgrasp_ref.so from the GRASP2018 sources where they lie
# under /root/reference (read-only), plus oracle/ref_driver.f90.  Test
# infrastructure only; nothing under /root/reference is copied.  Needs
# amdflang (ROCm).  Exits 0 without building when the reference tree or the
# compiler is absent.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${GRASP_REFERENCE:-/root/reference}/src"
FC="${FC:-/opt/rocm/bin/amdflang}"
OUT="$HERE/_ref"
if [ ! -d "$REF" ] || [ ! -x "$FC" ]; then
  echo "build_ref: reference tree or amdflang absent; skipping" >&2
  exit 0
fi
mkdir -p "$OUT/obj"
cd "$OUT/obj"
FFLAGS="-O2 -fPIC"
MODS="vast_kind_param_M parameter_def_M grid_C tatb_C wave_C orb_C def_C ncc_C cnc_C debug_C
      lic13_C sbc_C npot_C npar_C memory_man cteilsrk_C kkstart_C jlabl_C hmat_C
      bess_C bilst_C kkstartbreit_C wfac_C stor_C cons_C"
LIB9290="yzk_I quad_I dpbdt_I es_I slater_I dinit_I dmerge_I rinti_I
         radgrd setqic quad yzk slater rinti dpbdt nucpot es dinit dmerge spicmv2"
RCI="triangrk genintrk rkintc
     zkf_I rkint_I skint_I bessel_I brra_I brintf_I triangbreit1_I triangbreit2_I
     zkf rkint skint bessel brra brintf triangbreit1 triangbreit2 genintbreit1 genintbreit2
     brint1 brint2 brint3 brint4 brint5 brint6"
OBJS=""
for m in $MODS; do $FC $FFLAGS -c "$REF/lib/libmod/$m.f90" -o "$m.o"; OBJS="$OBJS $m.o"; done
for f in $LIB9290; do $FC $FFLAGS -c "$REF/lib/lib9290/$f.f90" -o "$f.o"; OBJS="$OBJS $f.o"; done
for f in $RCI; do $FC $FFLAGS -c "$REF/appl/rci90/$f.f90" -o "rci_$f.o"; OBJS="$OBJS rci_$f.o"; done
$FC $FFLAGS -c "$HERE/ref_driver.f90" -o ref_driver.o
$FC -shared -o "$OUT/libgrasp_ref.so" $OBJS ref_driver.o
rm -rf "$OUT/obj"
echo "build_ref: $OUT/libgrasp_ref.so"
# the angular-coefficient producer (librang90) behind the list fixtures
"$HERE/build_rang.sh"
