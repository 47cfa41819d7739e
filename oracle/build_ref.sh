#!/bin/bash
# Builds the reference md_run (Fortran 90 + OpenMP) from the sources where they
# lie under /root/reference/src_md_run, with ROCm flang, for use as a checker.
#
# Output: oracle/_ref/md_run only.  Intermediate files (the #SLCT/#SLC2 variant
# selections that the reference Makefile produces with its awk filter, objects
# and .mod files) are written to oracle/_ref/build and deleted at the end, so no
# copy of a reference source stays in the tree.
#
# The variant selection is the reference Makefile's own awk program
# (src_md_run/Makefile:126-132, 325-332, 368-374, 388-446) run by the POSIX awk
# in this image.  Floating-point contraction is switched off so the FP64 bits do
# not depend on FMA fusion (the production ifort build fuses).
set -euo pipefail
SRC=${JP_REF_SRC:-/root/reference/src_md_run}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
B=$OUT/build
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
FFLAGS="-O2 -cpp -fopenmp -ffp-contract=off -w"
[ -d "$SRC" ] || { echo "reference sources not found at $SRC"; exit 3; }
mkdir -p "$B"
cd "$B"

sel1() {  # sel1 SRC TAG OUT : keep #SLCT/#SLC2 lines carrying TAG
  awk -v t="#$2" '{ if ($0 ~ /#SLCT/) { if (index($0,t)) print $0 } else if ($0 ~ /#SLC2/) { if (index($0,t)) print $0 } else print $0 }' "$SRC/$1" > "$3"
}
selN() {  # selN SRC OUT TAG... : #SLCT lines need all tags, #SLC2 lines any tag
  local src=$1 out=$2; shift 2
  awk -v tags="$*" 'BEGIN{n=split(tags,T," ")}
    { if ($0 ~ /#SLCT/) { ok=1; for(i=1;i<=n;i++) if(!index($0,"#" T[i])) ok=0; if(ok) print $0 }
      else if ($0 ~ /#SLC2/) { ok=0; for(i=1;i<=n;i++) if(index($0,"#" T[i])) ok=1; if(ok) print $0 }
      else print $0 }' "$SRC/$src" > "$out"
}

MODS="COMPAR COMBAS COMERG COMMIS COMPSC COMCMM COMCMMC COMMOL PHYCNS CALC_TIME COMMPI"
for m in $MODS; do $FC $FFLAGS -c "$SRC/$m.f90"; done

for v in PB CB; do
  for f in cutoff dneres dnergv shake celpar; do sel1 $f.f90 $v ${f}_$v.f90; done
done
for e in CMM; do for p in REDU HIGH DOUB; do
  selN enear.f90 enear_${e}_$(echo $p | tr A-Z a-z).f90 $e $p; done; done
for e in ACT ZD; do for p in REDU HIGH DOUB; do for v in PB CB; do
  selN enear.f90 enear_${e}_$(echo $p | tr A-Z a-z)_$v.f90 $e $p $v; done; done; done

# flang rejects OPEN(STATUS='NEW') without FILE= at compile time (gfortran and
# ifort accept it).  flopen.f90 has three such branches, taken only when no file
# name is given, which md_run never does for the files used here.  The only
# change made to any reference source: those three STATUS='NEW' become
# STATUS='UNKNOWN'.
awk '{ if ($0 ~ /status=.new./ && $0 !~ /file=/) sub(/status=.new./,"status='"'"'unknown'"'"'"); print }' \
  "$SRC/flopen.f90" > flopen.f90
$FC $FFLAGS -c flopen.f90

PLAIN="main bstft calcm caldih calrsf caltmp chklin chkpha combin conj covang dnergy dnecap
 eigjac factor psr infjob inpbou inpcrb inpinp inpmd inpmin inpout inppdb inprst inptpb inptpl
 allocate_arrays input md mini mkabsr mksrdg monmd monmin outcpu outcpu2 outcrb outpdb output
 outrst outtpb ranno ranun dsr remotv rescrd scalet setshk setvel simulu sryelm staver
 steep verlet verlethe ellpbc ellprp inpcls rdlstatm setmnt inmcmd mnstmd inprep repprp celprp incel
 celoth celpol neartb cmmdo enear_extCMM setextCMM PBcord SIMD_check atom_specifier
 flclos efcol rdfree"
objs="flopen.o"
for f in $PLAIN; do $FC $FFLAGS -c "$SRC/$f.f90"; objs="$objs $f.o"; done
for f in inpvar dhr outtpl; do $FC $FFLAGS -c "$SRC/$f.f"; objs="$objs $f.o"; done
for f in *_PB.f90 *_CB.f90 enear_CMM_*.f90; do $FC $FFLAGS -c "$f"; objs="$objs ${f%.f90}.o"; done
modo=""; for m in $MODS; do modo="$modo $m.o"; done
$FC $FFLAGS $modo $objs -o "$OUT/md_run" -lm
cd "$OUT"
rm -rf "$B"
echo "built $OUT/md_run"
