#!/usr/bin/env bash
# Build the reference FORGE program (Fortran 90, /root/reference) for one
# compile-time configuration, plus the project-written harnesses
# oracle/ref_harness.f90 (calls the reference modules and dumps arrays) and
# oracle/ref_relax.f90 (the cell before and after LatticeRelaxation).
#
# The reference is configured only through `parameter`s in Setup.f90
# (Setup.f90:6-74); this recipe edits those values with sed in a scratch copy
# that lives under oracle/_ref/ only while the compiler runs and is deleted on
# exit.  Only the executables (and nothing else) stay in oracle/_ref/<case>/.
#
# Two edits beyond Setup parameters are needed because flang rejects constant
# out-of-bounds subscripts inside the dead numS=2 branches that ifort accepts:
#   Main.f90:58      nOccStates(numS), nPartOccStates(numS)  -> (2)
#   Main.f90:635-636 KineticEnergyIn(2), FockEnergyIn(2)     -> (numS)
# The first only enlarges arrays that the numS=1 path indexes at 1; the second
# only changes the never-executed numS=2 output branch.
# compile.sh:1's file order is also wrong (Geometry.f90:5 uses TightBinding),
# so the modules are compiled in dependency order here.
#
# Usage: [WRITEFOCK=1] [PRINTOUT=1] [C3=1] [NUMS=2] [ENFT=1] [ENFC2=1] [ENFC2T=1] [ENFV=1] oracle/ref_build.sh CASE NTHETA NUMK NFILLING ITMAX NREAD XI_NM [NLAYERS]
#   PRINTOUT=1 keeps the reference's nPrintOutputs = 1 (the Bands file, Main.f90:601-608).
#   C3=1 sets nenforceC3 = 1 (Setup.f90:27): Solve_C3 / GetFock_C3 replace Solve / GetFock.
#   NUMS=2 sets numS = numSIn = 2 (Setup.f90:13, 41): independent up/down spins.
#   ENFT / ENFC2 / ENFC2T = 1 set nenforceT / nenforceC2 / nenforceC2T (Setup.f90:28-30).
#   ENFV=1 sets nenforceValley = 1 (Setup.f90:31).
#   XI_NM is xi*0.246 (gate parameter in nm).  Toolchain: amdflang (ROCm) and
#   MKL LP64 sequential from /opt/conda/lib.
set -euo pipefail
CASE=$1; NTHETA=$2; NUMK=$3; NFILL=$4; ITMAX=$5; NREAD=$6; XINM=$7; NLAYERS=${8:-2}
WRITEFOCK=${WRITEFOCK:-0}
PRINTOUT=${PRINTOUT:-0}
C3=${C3:-0}
NUMS=${NUMS:-1}
ENFT=${ENFT:-0}
ENFC2=${ENFC2:-0}
ENFC2T=${ENFC2T:-0}
ENFV=${ENFV:-0}
REF=/root/reference
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref/$CASE
FC=${FC:-/opt/rocm/bin/amdflang}
MKL=${MKL_DIR:-/opt/conda/lib}
mkdir -p "$OUT"
SCR=$(mktemp -d "$HERE/_ref/.scratch.XXXXXX")
trap 'rm -rf "$SCR"' EXIT
for f in LapackRoutines Setup TightBinding HartreeFock Geometry Main; do
  cp "$REF/$f.f90" "$SCR/$f.f90"
done
if [ "$NLAYERS" = "3" ]; then ROT='[-1,+1,-1]'; else ROT='[-1,+1]'; fi
sed -i \
  -e "s/^integer(dp), parameter :: nlayers = .*/integer(dp), parameter :: nlayers = $NLAYERS/" \
  -e "s/^integer(dp), parameter :: ntheta = .*/integer(dp), parameter :: ntheta = $NTHETA/" \
  -e "s/^integer(dp), parameter :: RotateLayers(nlayers) = .*/integer(dp), parameter :: RotateLayers(nlayers) = $ROT/" \
  -e "s/^integer(dp), parameter :: numk = .*/integer(dp), parameter :: numk = $NUMK/" \
  -e "s/^integer(dp), parameter :: numkIn = .*/integer(dp), parameter :: numkIn = $NUMK/" \
  -e "s/^real(dp), parameter :: nfilling = .*/real(dp), parameter :: nfilling = ${NFILL}_dp/" \
  -e "s/^real(dp), parameter :: nfillingIn = .*/real(dp), parameter :: nfillingIn = ${NFILL}_dp/" \
  -e "s/^integer(dp), parameter :: itmax = .*/integer(dp), parameter :: itmax = $ITMAX/" \
  -e "s/^integer(dp), parameter :: nRead = .*/integer(dp), parameter :: nRead = $NREAD/" \
  -e "s/^integer(dp), parameter :: nPrintOutputs = .*/integer(dp), parameter :: nPrintOutputs = $PRINTOUT/" \
  -e "s/^integer(dp), parameter :: nWriteFock = .*/integer(dp), parameter :: nWriteFock = $WRITEFOCK/" \
  -e "s/^integer(dp), parameter:: nenforceC3 = .*/integer(dp), parameter:: nenforceC3 = $C3/" \
  -e "s/^integer(dp), parameter :: numS = .*/integer(dp), parameter :: numS = $NUMS/" \
  -e "s/^integer(dp), parameter:: nenforceT = .*/integer(dp), parameter:: nenforceT = $ENFT/" \
  -e "s/^integer(dp), parameter:: nenforceC2 = .*/integer(dp), parameter:: nenforceC2 = $ENFC2/" \
  -e "s/^integer(dp), parameter:: nenforceC2T = .*/integer(dp), parameter:: nenforceC2T = $ENFC2T/" \
  -e "s/^integer(dp), parameter:: nenforceValley = .*/integer(dp), parameter:: nenforceValley = $ENFV/" \
  -e "s/^integer(dp), parameter :: numSIn = .*/integer(dp), parameter :: numSIn = $NUMS/" \
  -e "s/^real(dp), parameter:: xi = .*/real(dp), parameter:: xi = ${XINM}_dp\/0.246_dp/" \
  -e "s/^real(dp), parameter:: xiIn = .*/real(dp), parameter:: xiIn = ${XINM}_dp\/0.246_dp/" \
  "$SCR/Setup.f90"
grep -q "ntheta = $NTHETA\$" "$SCR/Setup.f90"
grep -q "numk = $NUMK\$" "$SCR/Setup.f90"
grep -q "nenforceC3 = $C3\$" "$SCR/Setup.f90"
grep -q "numS = $NUMS\$" "$SCR/Setup.f90"
grep -q "nenforceT = $ENFT\$" "$SCR/Setup.f90"
grep -q "nenforceC2 = $ENFC2\$" "$SCR/Setup.f90"
grep -q "nenforceC2T = $ENFC2T\$" "$SCR/Setup.f90"
grep -q "nenforceValley = $ENFV\$" "$SCR/Setup.f90"
sed -i \
  -e '58s/nOccStates(numS), nPartOccStates(numS)/nOccStates(2), nPartOccStates(2)/' \
  -e '635s/KineticEnergyIn(2)/KineticEnergyIn(numS)/' \
  -e '636s/FockEnergyIn(2)/FockEnergyIn(numS)/' \
  "$SCR/Main.f90"
cp "$HERE/ref_harness.f90" "$SCR/ref_harness.f90"
cp "$HERE/ref_relax.f90" "$SCR/ref_relax.f90"
cd "$SCR"
FFLAGS="-O2 -fopenmp -w"
for f in LapackRoutines Setup TightBinding HartreeFock Geometry; do
  $FC $FFLAGS -c $f.f90
done
LIBS="-L$MKL -Wl,-rpath,$MKL -lmkl_gf_lp64 -lmkl_sequential -lmkl_core -lpthread -lm -ldl"
$FC $FFLAGS -o "$OUT/forge.x" LapackRoutines.o Setup.o TightBinding.o HartreeFock.o Geometry.o Main.f90 $LIBS
# the harness dumps the spin-singlet arrays only
if [ "$NUMS" = "1" ]; then
  $FC $FFLAGS -o "$OUT/harness.x" LapackRoutines.o Setup.o TightBinding.o HartreeFock.o Geometry.o ref_harness.f90 $LIBS
  $FC $FFLAGS -o "$OUT/relax.x" LapackRoutines.o Setup.o TightBinding.o HartreeFock.o Geometry.o ref_relax.f90 $LIBS
fi
echo "built $OUT (ntheta=$NTHETA numk=$NUMK nfilling=$NFILL itmax=$ITMAX nRead=$NREAD xi=${XINM}nm nlayers=$NLAYERS)"
