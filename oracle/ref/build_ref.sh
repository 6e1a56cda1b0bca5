#!/usr/bin/env bash
# Build the reference SuLi-IC solver routines with LLVM flang, for one grid
# configuration, together with the harness program in this directory.
#
# The reference fixes its grid, boundary codes and model switches as Fortran
# `parameter`s (0_modules.f90), so every configuration is its own build.  This
# script reads the sources where they lie (/root/reference/SuLi-IC), writes a
# configured copy into oracle/_ref/<tag>/src (ignored by git, never committed)
# and compiles it to oracle/_ref/<tag>/harness.  The edits are configuration
# only, plus one rename that LLVM flang needs:
#   0_modules.f90:7      dx, dy, dz
#   0_modules.f90:15     nx, ny, nz
#   0_modules.f90:20     uinicial = 0 (fluid at rest)
#   0_modules.f90:24     t_tempo (optional 18th argument, default 1: RK2)
#   0_modules.f90:26-27  der, adv_type (optional 19th and 20th, default 1, 1)
#   0_modules.f90:29     obst_t (optional 21st, default 0; the harness loads ku, kv, kw)
#   0_modules.f90:31     m_turb
#   0_modules.f90:53-61  boundary codes ccx0..cczf
#   0_modules.f90:143    gx, gz
#   0_modules.f90:181    csmag
#   module `tempo` (0_modules.f90:153) -> `tempo_mod`: flang rejects
#   `use tempo` inside 11_diversos.f90, whose SUBROUTINE tempo has the same
#   name (11_diversos.f90:10,237).
# 1_principal.f90 (the PROGRAM) is left out; harness.f90 replaces it and
# calls the reference subroutines in the order of 1_principal.f90:64-108.
#
# REF_VARIANT (environment) selects a CPU-baseline build, with the tag
# suffix _<variant>:
#   bench      -O3 as the reference's Makefile:7, and graddin prints its CG
#              iteration count and final residual after its loop
#              (9_grad.f90:267, a write statement; nothing else changes)
#   bench_rel  as bench, with the stopping rule of 9_grad.f90:175 replaced by
#              the relative one sqrt(sum r_s^2) <= 1e-8 ||b_s|| (b_s = q/sqrt(D),
#              the scaled right-hand side; computed once before the loop)
# Without it the build is -O2 (the parity harnesses).
#
# REF_MMS="MMS_T A H0 COEF" (environment) builds the manufactured-solution
# configuration: 0_modules.f90:37 mms_t and the module mms_m constants
# a, h0, coef (0_modules.f90:208-212).
# REF_WAVE=WAVE_T (environment) builds the wave inflow: 0_modules.f90:35.
#
# usage: build_ref.sh TAG NX NY NZ DX DY DZ CCX0 CCXF CCY0 CCYF CCZ0 CCZF MTURB CSMAG GX GZ [TTEMPO [DER ADV_TYPE [OBST_T]]]
set -euo pipefail
if [ $# -ne 17 ] && [ $# -ne 18 ] && [ $# -ne 20 ] && [ $# -ne 21 ]; then
  echo "usage: $0 TAG NX NY NZ DX DY DZ CCX0 CCXF CCY0 CCYF CCZ0 CCZF MTURB CSMAG GX GZ [TTEMPO [DER ADV_TYPE [OBST_T]]]" >&2
  exit 2
fi
TAG=$1; NX=$2; NY=$3; NZ=$4; DX=$5; DY=$6; DZ=$7
CCX0=$8; CCXF=$9; CCY0=${10}; CCYF=${11}; CCZ0=${12}; CCZF=${13}
MTURB=${14}; CSMAG=${15}; GX=${16}; GZ=${17}; TTEMPO=${18:-1}; DER=${19:-1}; ADVT=${20:-1}; OBST=${21:-0}

HERE=$(cd "$(dirname "$0")" && pwd)
REF=${SULI_REFERENCE:-/root/reference}/SuLi-IC
OUT=$HERE/../_ref/$TAG
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
if [ ! -d "$REF" ]; then echo "reference sources not found at $REF" >&2; exit 3; fi

rm -rf "$OUT"
mkdir -p "$OUT/src" "$OUT/obj"
for f in 0_modules 2_iniciais 3_contorno 4_plot 5_viscosidade 6_convdiff 7_advectivo \
         8_derivadas 9_grad 10_pos 11_diversos 12_paraview_3D 13_mms 14_level_set; do
  cp "$REF/$f.f90" "$OUT/src/$f.f90"
done
chmod u+w "$OUT"/src/*.f90

M=$OUT/src/0_modules.f90
edit() {  # edit FILE SED-EXPR PATTERN-THAT-MUST-EXIST-AFTER
  sed -i -e "$2" "$1"
  grep -q -- "$3" "$1" || { echo "edit failed in $1: $2" >&2; exit 4; }
}
edit "$M" "s/^real(8),parameter :: dx = 0.02, dy = 0.02, dz = 0.02/real(8),parameter :: dx = $DX, dy = $DY, dz = $DZ/" "dx = $DX, dy = $DY, dz = $DZ"
edit "$M" "s/^integer,parameter :: nx=int(.5\/dx) , ny=int(1.\/dy), nz=int(0.9\/dz)/integer,parameter :: nx=$NX, ny=$NY, nz=$NZ/" "nx=$NX, ny=$NY, nz=$NZ"
edit "$M" "s/^real(8),parameter :: uinicial = 0.31/real(8),parameter :: uinicial = 0./" "uinicial = 0\."
edit "$M" "s/^integer,parameter :: t_tempo = 1 /integer,parameter :: t_tempo = $TTEMPO /" "t_tempo = $TTEMPO "
edit "$M" "s/^integer,parameter :: der = 1 /integer,parameter :: der = $DER /" "der = $DER "
edit "$M" "s/^integer,parameter :: adv_type = 1 /integer,parameter :: adv_type = $ADVT /" "adv_type = $ADVT "
edit "$M" "s/^integer,parameter :: obst_t = 0 /integer,parameter :: obst_t = $OBST /" "obst_t = $OBST "
edit "$M" "s/^integer,parameter :: m_turb = 0/integer,parameter :: m_turb = $MTURB/" "m_turb = $MTURB"
edit "$M" "s/^integer,parameter :: ccx0=0/integer,parameter :: ccx0=$CCX0/" "ccx0=$CCX0"
edit "$M" "s/^integer,parameter :: ccxf=0/integer,parameter :: ccxf=$CCXF/" "ccxf=$CCXF"
edit "$M" "s/^integer,parameter :: ccy0=2/integer,parameter :: ccy0=$CCY0/" "ccy0=$CCY0"
edit "$M" "s/^integer,parameter :: ccyf=2/integer,parameter :: ccyf=$CCYF/" "ccyf=$CCYF"
edit "$M" "s/^integer,parameter :: ccz0=2/integer,parameter :: ccz0=$CCZ0/" "ccz0=$CCZ0"
edit "$M" "s/^integer,parameter :: cczf=1/integer,parameter :: cczf=$CCZF/" "cczf=$CCZF"
edit "$M" "s/^real(8), parameter :: gx = 9.80665 \* sin(atan(decliv)) , gz = 9.80665 \* cos(atan(decliv))/real(8), parameter :: gx = $GX , gz = $GZ/" "gx = $GX , gz = $GZ"
edit "$M" "s/^real(8), parameter :: csmag=0.13/real(8), parameter :: csmag=$CSMAG/" "csmag=$CSMAG"
if [ -n "${REF_MMS:-}" ]; then
  read -r MMST MMSA MMSH0 MMSC <<<"$REF_MMS"
  edit "$M" "s/^integer,parameter :: mms_t = 0 /integer,parameter :: mms_t = $MMST /" "mms_t = $MMST "
  edit "$M" "s/^real(8), parameter :: a  = 0\.!5\$/real(8), parameter :: a  = $MMSA/" "a  = $MMSA\$"
  edit "$M" "s/^real(8), parameter :: h0 = 2\.\$/real(8), parameter :: h0 = $MMSH0/" "h0 = $MMSH0\$"
  edit "$M" "s/^real(8), parameter :: coef = 0\. ! 1\.\$/real(8), parameter :: coef = $MMSC/" "coef = $MMSC\$"
fi
if [ -n "${REF_WAVE:-}" ]; then
  edit "$M" "s/^integer,parameter :: wave_t = 0 /integer,parameter :: wave_t = $REF_WAVE /" "wave_t = $REF_WAVE "
fi
edit "$M" "s/^module tempo\$/module tempo_mod/" "^module tempo_mod"
edit "$M" "s/^end module tempo\$/end module tempo_mod/" "^end module tempo_mod"
for f in 2_iniciais 4_plot 5_viscosidade 11_diversos; do
  sed -i -e 's/^\([[:space:]]*\)[Uu][Ss][Ee] tempo[[:space:]]*$/\1USE tempo_mod/' "$OUT/src/$f.f90"
done
if grep -n -i "^[[:space:]]*use tempo[[:space:]]*$" "$OUT"/src/*.f90; then
  echo "unrenamed 'use tempo' left" >&2; exit 4
fi

OPT=-O2
VARIANT=${REF_VARIANT:-}
if [ -n "$VARIANT" ]; then
  G=$OUT/src/9_grad.f90
  OPT=-O3
  case "$VARIANT" in
    bench|bench_rel) ;;
    *) echo "unknown REF_VARIANT $VARIANT" >&2; exit 2 ;;
  esac
  edit "$G" "/^[[:space:]]*! Desnormaliz/i\\		write(*,'(A,I0,A,ES16.8)') 'REF_CG_ITERS ', cont, ' res2=', abs(alfamupr)" "REF_CG_ITERS"
  if [ "$VARIANT" = bench_rel ]; then
    edit "$G" "s/^\\([[:space:]]*\\)real(8) :: aux1, aux2, aux3\$/\\1real(8) :: aux1, aux2, aux3, tolrel2/" "aux3, tolrel2"
    edit "$G" "/^[[:space:]]*do while ((abs(alfamupr) > (0.0001\\/(nx\\*ny\\*nz)))/i\\		tolrel2 = 1e-16 * sum((matqpr/sqrt(matdpr(1:nx,1:ny,1:nz)))**2)" "tolrel2 = 1e-16"
    edit "$G" "s/(abs(alfamupr) > (0.0001\\/(nx\\*ny\\*nz)))/(abs(alfamupr) > tolrel2)/" "(abs(alfamupr) > tolrel2)"
  fi
fi
FFLAGS="$OPT -fdefault-real-8 -fdefault-double-8 -mcmodel=medium"
cd "$OUT/obj"
for f in 0_modules 2_iniciais 3_contorno 4_plot 5_viscosidade 6_convdiff 7_advectivo \
         8_derivadas 9_grad 10_pos 11_diversos 12_paraview_3D 13_mms 14_level_set; do
  "$FC" $FFLAGS -c "$OUT/src/$f.f90" -o "$f.o"
done
"$FC" $FFLAGS -c "$HERE/harness.f90" -o harness.o
"$FC" $OPT -mcmodel=medium -o "$OUT/harness" ./*.o
echo "built $OUT/harness (nx=$NX ny=$NY nz=$NZ${VARIANT:+, variant $VARIANT})"
