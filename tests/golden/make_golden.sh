#!/bin/bash
# Regenerates the golden fixtures under tests/golden/ (run in the build
# container, where the reference sources exist):
#  1. data files the reference's own tests hold (fixture CSVs, the elevate
#     PLOT3D blade grid), copied as data;
#  2. CL/CT histories of the reference solver (oracle/_ref/volcanor, compiled
#     from its own sources by oracle/Makefile) on our case directories.
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
REF=/root/reference/tests
cp "$REF/katzNplotkin-AR04.case/referenceResults/r01ForceNonDim.csv.ref" "$HERE/katz_plotkin_AR4_ForceNonDim.csv"
cp "$REF/katzNplotkin-AR04.case/referenceResults/r01b01ForceDist00160.csv.ref" "$HERE/katz_plotkin_AR4_ForceDist00160.csv"
cp "$REF/elevateTest.case/referenceResults/r01ForceNonDim.csv.ref" "$HERE/elevate_ForceNonDim.csv"
cp "$REF/elevateTest.case/referenceResults/r01b01ForceDist00150.csv.ref" "$HERE/elevate_ForceDist00150.csv"
cp "$REF/katzNplotkin-AR04.case/referenceResults/r01Params.json.ref" "$HERE/katz_plotkin_AR4_Params.json"
cp "$REF/elevateTest.case/referenceResults/r01Params.json.ref" "$HERE/elevate_Params.json"
cp "$REF/elevateTest.case/elevate.xyz" "$HERE/elevate.xyz"
make -C "$ROOT/oracle" ref
python3 "$ROOT/cases/make_cases.py"
run_ref() {  # case-name  output-name
  local tmp; tmp=$(mktemp -d)
  cp "$ROOT/cases/$1.case/"* "$tmp/"
  mkdir -p "$tmp/Results" "$tmp/Restart"
  (cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 OMP_SCHEDULE=static \
     "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
  cp "$tmp/Results/r01ForceNonDim.csv" "$HERE/$2"
  for f in "$tmp"/Results/r0[2-9]ForceNonDim.csv; do
    [ -e "$f" ] && cp "$f" "$HERE/${2%_ForceNonDim.csv}_$(basename "$f")"
  done
  rm -rf "$tmp"
}
[ -e "$HERE/cfg1_wing_ForceNonDim.csv" ] || run_ref cfg1_wing cfg1_wing_ForceNonDim.csv
run_ref cfg2_rotor_2rev cfg2_rotor_2rev_ForceNonDim.csv
run_ref cfg2_rotor_farwake_2rev cfg2_rotor_farwake_2rev_ForceNonDim.csv
# 3. per-step wing and wake state of the reference (oracle/_ref/refprobe,
#    our driver linked against the reference objects) on the small
#    forward-flight rotor: ring corners, core radii, circulation, velocity
#    histories after the initial solution and after each of 14 steps.
tmp=$(mktemp -d)
cp "$ROOT/cases/refprobe_rotor.case/"* "$tmp/"
mkdir -p "$tmp/Results"
(cd "$tmp" && OMP_NUM_THREADS=1 "$ROOT/oracle/_ref/refprobe" > run.log 2>&1)
cp "$tmp/Results/refprobe.txt" "$HERE/refprobe_rotor_wake.txt"
rm -rf "$tmp"
# 4. multi-rotor coupling: quad-rotor (one revolution, far wake and
#    truncation within it) and rotor above a wing (half a revolution)
run_ref cfg4_quad_short cfg4_quad_short_ForceNonDim.csv
run_ref cfg5_rotor_wing_short cfg5_rotor_wing_short_ForceNonDim.csv
# 5. the FT1, AB2, AM3 and AM4 wake schemes on a wing and a forward-flight rotor
for tag in ft1 ab2 am3 am4; do
  run_ref scheme_${tag}_wing scheme_${tag}_wing_ForceNonDim.csv
  run_ref scheme_${tag}_rotor scheme_${tag}_rotor_ForceNonDim.csv
done
# 5b. the start-up wake velocity initWakeVel under AM2 and AM4
for tag in am2 am4; do
  run_ref initwake_${tag}_rotor initwake_${tag}_rotor_ForceNonDim.csv
done
# 5c. slow start and sub-iterations
for k in 1 2 3; do
  run_ref slowstart${k}_rotor slowstart${k}_rotor_ForceNonDim.csv
done
run_ref subiter_rotor_wing subiter_rotor_wing_ForceNonDim.csv
run_ref burst_rotor burst_rotor_ForceNonDim.csv
run_ref traj_wing traj_wing_ForceNonDim.csv
run_ref camber_wing camber_wing_ForceNonDim.csv
run_ref restart_rotor restart_rotor_ForceNonDim.csv
# the rotor above the wing under AM2, with the restart write switched off:
# the reference's restart write aborts for every scheme but AM4
tmp=$(mktemp -d)
cp "$ROOT/cases/restart_rotor_wing.case/"* "$tmp/"
sed -i 's/restartWriteNt = 18/restartWriteNt = 0/' "$tmp/config.nml"
mkdir -p "$tmp/Results" "$tmp/Restart"
(cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 OMP_SCHEDULE=static \
   "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
cp "$tmp/Results/r01ForceNonDim.csv" "$HERE/restart_rotor_wing_ForceNonDim.csv"
cp "$tmp/Results/r02ForceNonDim.csv" "$HERE/restart_rotor_wing_r02ForceNonDim.csv"
rm -rf "$tmp"
# 5d. wing and wake plots (.tec) of a forward-flight rotor at steps 20 and 40, and
#     the force files of every force step (ForceNonDim, ForceDim, ForceDist)
tmp=$(mktemp -d)
cp "$ROOT/cases/tec_rotor.case/"* "$tmp/"
mkdir -p "$tmp/Results" "$tmp/Restart"
(cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 OMP_SCHEDULE=static \
   "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
for st in 00020 00040; do
  cp "$tmp/Results/r01wingNwake$st.tec" "$HERE/tec_rotor_wingNwake$st.tec"
done
for f in ForceNonDim.csv ForceDim.csv b01ForceDist00000.csv b02ForceDist00001.csv \
         b01ForceDist00040.csv b02ForceDist00040.csv bladedynamics.csv b01Surface.tec \
         b02Surface.tec; do
  cp "$tmp/Results/r01$f" "$HERE/tec_rotor_$f"
done
rm -rf "$tmp"
# 5e. inflow and bound-circulation plots of a rotor above a wing
tmp=$(mktemp -d)
cp "$ROOT/cases/plots_rotor_wing.case/"* "$tmp/"
mkdir -p "$tmp/Results" "$tmp/Restart"
(cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 OMP_SCHEDULE=static \
   "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
for f in "$tmp"/Results/*.curve; do
  cp "$f" "$HERE/plots_rotor_wing_$(basename "$f")"
done
rm -rf "$tmp"
# 5f. the velocity field on a grid: the reference's gridgen on the reference's
#     filament dumps of the grid_rotor case (dump sizes are in the test)
tmp=$(mktemp -d)
cp "$ROOT/cases/grid_rotor.case/"* "$tmp/"
mkdir -p "$tmp/Results" "$tmp/Restart"
(cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 OMP_SCHEDULE=static \
   "$ROOT/oracle/_ref/volcanor" > run.log 2>&1 && "$ROOT/oracle/_ref/gridgen" > gridgen.log 2>&1)
for st in 00020 00040; do
  cp "$tmp/Results/grid$st.tec" "$HERE/grid_rotor_grid$st.tec"
done
rm -rf "$tmp"
# 5g. the cambered wing's blade surface file
tmp=$(mktemp -d)
cp "$ROOT/cases/camber_wing.case/"* "$tmp/"
mkdir -p "$tmp/Results" "$tmp/Restart"
(cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
cp "$tmp/Results/r01b01Surface.tec" "$HERE/camber_wing_b01Surface.tec"
rm -rf "$tmp"
# 5h. near-wake skew plots of a forward-flight rotor
tmp=$(mktemp -d)
cp "$ROOT/cases/skew_rotor.case/"* "$tmp/"
mkdir -p "$tmp/Results" "$tmp/Restart"
(cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
for f in "$tmp"/Results/*skew*.csv; do
  cp "$f" "$HERE/skew_rotor_$(basename "$f")"
done
rm -rf "$tmp"
# 5h. wake tip plots of a forward-flight rotor in slow start, both directions
for c in tip_rotor tip_rotor_cw; do
  tmp=$(mktemp -d)
  cp "$ROOT/cases/$c.case/"* "$tmp/"
  mkdir -p "$tmp/Results" "$tmp/Restart"
  (cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
  cp "$tmp/Results/r01ForceNonDim.csv" "$HERE/${c}_ForceNonDim.csv"
  for f in "$tmp"/Results/r01tip*; do
    cp "$f" "$HERE/${c}_$(basename "$f")"
  done
  rm -rf "$tmp"
done
# 5i. prescribed far wake: CL/CT history and wing and wake plots at steps 24 and 48
for c in presc_rotor presc_gen_rotor; do
  tmp=$(mktemp -d)
  cp "$ROOT/cases/$c.case/"* "$tmp/"
  mkdir -p "$tmp/Results" "$tmp/Restart"
  (cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 OMP_SCHEDULE=static \
     "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
  cp "$tmp/Results/r01ForceNonDim.csv" "$HERE/${c}_ForceNonDim.csv"
  for st in 00024 00048; do
    cp "$tmp/Results/r01wingNwake$st.tec" "$HERE/${c}_wingNwake$st.tec"
  done
  rm -rf "$tmp"
done
# 5j. image rotors: CL/CT histories of every rotor, wing and wake plots at
# steps 24 and 48, and (image_rotor_xy) tip plots at steps 24 and 48
for c in image_rotor image_rotor_am4 image_rotor_xy; do
  tmp=$(mktemp -d)
  cp "$ROOT/cases/$c.case/"* "$tmp/"
  mkdir -p "$tmp/Results" "$tmp/Restart"
  (cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 OMP_SCHEDULE=static \
     "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
  for r in r01 r02 r03; do
    [ -f "$tmp/Results/${r}ForceNonDim.csv" ] || continue
    cp "$tmp/Results/${r}ForceNonDim.csv" "$HERE/${c}_${r}ForceNonDim.csv"
    for st in 00024 00048; do
      cp "$tmp/Results/${r}wingNwake$st.tec" "$HERE/${c}_${r}wingNwake$st.tec"
      [ -f "$tmp/Results/${r}tip$st.tec" ] && cp "$tmp/Results/${r}tip$st.tec" "$HERE/${c}_${r}tip$st.tec"
    done
  done
  rm -rf "$tmp"
done
# 6. velocity probes: the reference's probes2file output of a short wing run
tmp=$(mktemp -d)
cp "$ROOT/cases/probes_wing.case/"* "$tmp/"
mkdir -p "$tmp/Results" "$tmp/Restart"
(cd "$tmp" && ulimit -s unlimited && OMP_NUM_THREADS=8 "$ROOT/oracle/_ref/volcanor" > run.log 2>&1)
for s in 00001 00010 00020; do cp "$tmp/Results/probes$s.csv" "$HERE/probes_wing_probes$s.csv"; done
rm -rf "$tmp"
