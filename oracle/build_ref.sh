#!/usr/bin/env bash
# Build recipe for oracle/_ref/libref.so: the reference's own Fortran sources
# (sscollis/lns3d src/) compiled with AMD flang, plus this repo's direct-call
# harness oracle/refdrv.f90.  Test infrastructure only: nothing in the product
# links or loads it.
#
# * Compiled from the sources where they lie (REF=/root/reference by default);
#   objects go to a temporary directory that is removed afterwards, and the
#   only output is oracle/_ref/libref.so.
# * Flags follow src/gcc.mak:18-20 (-fdefault-real-8 -fdefault-double-8, -cpp,
#   -DUSE_TRANSIENT_EIGENFUNCTION, 120-column fixed form), with -fPIC added.
# * One deviation: src/cpenta1bc.f:675 and src/cpenta2bc.f:676 index the rhs
#   array in the pre-2000 (nsys,n,nblk) order; they are compiled from a
#   temporary copy with that single line corrected to the (nblk,nsys,n) /
#   (nblk,n,nsys) order used by every other line of the same files.  The
#   shipped routines are also compiled unchanged, with their symbols renamed
#   to cpenta1bc_shipped_ / cpenta2bc_shipped_, so tests can show the defect.
# * The main program (src/lns3d.f90) is left out; the harness calls routines
#   directly.
# * Second output, oracle/_ref/libref_hip.so: the harness relinked against the
#   HIP library's Fortran drop-ins (below).
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF="${LNS3D_REFERENCE:-/root/reference}"
SRC="$REF/src"
OUT="$HERE/_ref"
if [ ! -d "$SRC" ]; then
  echo "build_ref.sh: reference sources not found at $SRC; skipping" >&2
  exit 0
fi
FC="${FC:-flang}"
command -v "$FC" >/dev/null || FC=/opt/rocm/lib/llvm/bin/flang
COMMON="-cpp -fdefault-real-8 -fdefault-double-8 -DUSE_TRANSIENT_EIGENFUNCTION -O2 -fopenmp -fPIC"
F90FLAGS="$COMMON"
F77FLAGS="$COMMON -ffixed-line-length=120"

TMP="$(mktemp -d /tmp/lns3d_ref.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT
cd "$TMP"

MODS="global stencil local2d buffer sponge local3d pot bump temporal"
OBJS="lhs penta1p lhsbc1 penta2bc checkbc gendist lhsbc1f penta2p damper gengrid
lhsbc2 resid genini lhsbc2f resstat lhsbt1 genmean lhsbs1f error genmtrx lhsbt1f
expdrv lhsbt2f bcfix calcp filter lhsf rhsbc fstat2d getmat g1 lhsbt2 g2 grad
lrhs smoother grad2 rk rkbc gradbc nrk sstat nsolve2p impdrv nsolve1bc lhsbs1
nsolve1p traces input nsolve2bc itrbc penta1bc wdata reimann genini3d cgrad
impdrv3d resstat3d cgrad2 cgradbc itrbc3d lrhs3d rhsbc3d cpenta1p cpenta2p
init lhs1f3d lhsbt1f3d lhsbt2f3d lhsbc1f3d lhsbc2f3d lhsbs1f3d expdrv3d rk3d
rkbc3d lhs2f3d fstat3d potential genbump smoother3d misc rhs_l rpenta1p
rpenta2p rpenta1bc rpenta2bc lwallbc wallbc gena_n genb_n gena_l genb_l second"

compile() {  # $1 = basename, $2 = source path
  case "$2" in
    *.f)   "$FC" $F77FLAGS -c "$2" -o "$1.o" ;;
    *.f90) "$FC" $F90FLAGS -c "$2" -o "$1.o" ;;
  esac
}
for m in $MODS; do compile "$m" "$SRC/$m.f90"; done
for o in $OBJS; do
  if [ -f "$SRC/$o.f90" ]; then compile "$o" "$SRC/$o.f90";
  elif [ -f "$SRC/$o.f" ]; then compile "$o" "$SRC/$o.f";
  else echo "missing $o" >&2; exit 1; fi
done
gcc -O -fPIC -c "$SRC/getver.c" -o getver.o

# corrected copies of the two non-periodic complex solvers (one line each)
sed '675s/rhs(iv,r,m) = rhs(iv,r,m) + mult \* rhs(iv,i,nc)/rhs(m,iv,r) = rhs(m,iv,r) + mult * rhs(nc,iv,i)/' \
    "$SRC/cpenta1bc.f" > cpenta1bc_fixed.f
sed '676s/rhs(iv,m,r) = rhs(iv,m,r) + mult \* rhs(iv,nc,i)/rhs(m,r,iv) = rhs(m,r,iv) + mult * rhs(nc,i,iv)/' \
    "$SRC/cpenta2bc.f" > cpenta2bc_fixed.f
if cmp -s cpenta1bc_fixed.f "$SRC/cpenta1bc.f" || cmp -s cpenta2bc_fixed.f "$SRC/cpenta2bc.f"; then
  echo "build_ref.sh: the index correction did not apply" >&2; exit 1
fi
compile cpenta1bc cpenta1bc_fixed.f
compile cpenta2bc cpenta2bc_fixed.f
# shipped versions, renamed
compile cpenta1bc_shipped "$SRC/cpenta1bc.f"
compile cpenta2bc_shipped "$SRC/cpenta2bc.f"
objcopy --redefine-sym cpenta1bc_=cpenta1bc_shipped_ cpenta1bc_shipped.o
objcopy --redefine-sym cpenta2bc_=cpenta2bc_shipped_ cpenta2bc_shipped.o

"$FC" $F90FLAGS -I. -c "$HERE/refdrv.f90" -o refdrv.o

mkdir -p "$OUT"
"$FC" -shared -fopenmp -o "$OUT/libref.so.tmp" *.o
mv "$OUT/libref.so.tmp" "$OUT/libref.so"
echo "built $OUT/libref.so"

# The same harness with the reference's lrhs3D, lhs1f3D, lhs2f3D, cgrad,
# cgrad2 and cpenta* objects left out, so that those names resolve to the
# HIP library's Fortran drop-ins (lns3d_amd/liblns3d_fortran.so, whose module
# file lns3d_hip.mod is in lns3d_amd/_build), plus the glue that forwards the
# module state (refdrv_hip.f90): oracle/_ref/libref_hip.so.  Built only where
# the HIP library has been built.
PKG="$HERE/../lns3d_amd"
if [ -f "$PKG/liblns3d_fortran.so" ] && [ -f "$PKG/_build/lns3d_hip.mod" ]; then
  mkdir hip
  "$FC" $F90FLAGS -I. -I"$PKG/_build" -c "$HERE/refdrv_hip.f90" -o hip/refdrv_hip.o
  KEEP=""
  for o in *.o; do
    case "$o" in
      lrhs3d.o|lhs1f3d.o|lhs2f3d.o|cgrad.o|cgrad2.o|cpenta1bc.o|cpenta2bc.o|cpenta1p.o|cpenta2p.o) ;;
      *) KEEP="$KEEP $o" ;;
    esac
  done
  "$FC" -shared -fopenmp -o "$OUT/libref_hip.so.tmp" $KEEP hip/refdrv_hip.o \
      -L"$PKG" -llns3d_fortran -llns3d_amd -Wl,-rpath,'$ORIGIN/../../lns3d_amd'
  mv "$OUT/libref_hip.so.tmp" "$OUT/libref_hip.so"
  echo "built $OUT/libref_hip.so"
fi
