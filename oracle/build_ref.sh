#!/usr/bin/env bash
# Builds oracle/_ref/libyetiref.so from the YETI Fortran sources where they lie (default
# /root/reference/src/yeti_iga), plus oracle/ref_shim.f90. Test infrastructure only.
#
# The file list is the reference's own `base` library (src/yeti_iga/CMakeLists.txt:21-58),
# the `optim` library (src/yeti_iga/optim/CMakeLists.txt:20-27), the two assembly drivers,
# the displacement reconstruction and evaluation, the mass-matrix driver, the VTU writer
# (generateVTU with its per-element-type field routines and writePatch), the compliance
# gradient driver and the natural-frequency gradient (gradVibration_AN, gradUELMASSMAT1),
# compiled with the reference's flags (-O2, -fopenmp, -fPIC; CMakeLists.txt:69,
# src/CMakeLists.txt:42-48) by amdflang (LLVM flang), the only Fortran compiler in this image.
#
# One compatibility rewrite is applied to a temporary copy (never kept): flang rejects the
# specification expression `indDLoad(SUM(load_target_nbelem))` because load_target_nbelem is
# given its dimension later in the same statement (gfortran accepts it). The bound is
# rewritten to assumed size, `indDLoad(*)`; the routines only index indDLoad inside that
# range, so behaviour is unchanged. Nothing else is modified. Only the .so is written into
# oracle/_ref/; the temporary sources and objects are deleted.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${YETI_REF_SRC:-/root/reference/src/yeti_iga}"
OUT="$HERE/_ref"
FC="${FC:-amdflang}"
if [ ! -d "$REF" ]; then
    echo "build_ref: $REF not present, skipping oracle/_ref" >&2
    exit 0
fi
FILES="parameters.f90 igaparmetrization.f90 embeddedmapping.f90 main/material_lib.f90
main/material_lib_HO.f90 main/dersbasisfuns.f90 main/shap.f90 main/shap_HO.f90 main/shapPress.f
main/operateurs.f main/GaussLegendre.f90 main/UELMAT.f90 main/UELMAT_HO.f90 main/stiffmatrix.F90
main/stiffmatrix_HO.f90 main/UMASSMAT.f main/evaluatenurbsfcts.f main/evaluateStress.f
main/point_inversion.f90 main/projection.f90 main/isElemOnFace.f90 main/shell/nurbsbasisfuns.f
main/shell/USFMEM_shell.f main/shell/USFBND_shell.f main/shell/curvilinearCoordinates.f
main/shell/shapPress_shell.f main/shell/UELMAT_shell.f main/shell/ComputeMemStrain_shell.f
main/shell/ComputeBndStrain_shell.f main/shell/evaluateStress_shell.f
main/embedded/mapEntities2parametricSpaces.f main/embedded/UELMAT_embeddedshell.f
main/embedded/UELMAT_embeddedvol.f90 main/m_nurbspatch.f90 main/shapPress_omp.f90
main/shap_omp.f90 main/uelmat_solid_omp.f90 main/build_system_elemStorage.f90
main/build_system_elemStorage_OMP.f90 main/reconstruction.f postprocessing/evaluate_disp.f
main/compute_CMASSMatrix.f postprocessing/compute_svars_solid.f postprocessing/compute_svars_shell.f
postprocessing/compute_svars_embdedshell.f postprocessing/compute_svars_embdedsolid.f
postprocessing/writepatch.f postprocessing/postproc_vtu.f optim/gradUELMAT_solid.f90 optim/gradUELMAT_shell.f
optim/gradUELMAT_embeddedshell.f optim/gradUELMAT_embeddedsolid.f90 optim/gradtotalwork_fullAN.f90
optim/gradUELMASSMAT_solid.f optim/gradvibration_fullAN.f"
TMP="$(mktemp -d /tmp/yetiref.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT
cd "$TMP"
for f in $FILES; do
    src="$REF/$f"
    b="$(basename "$f")"
    if grep -q 'indDLoad(SUM(load_target_nbelem))' "$src"; then
        sed 's/indDLoad(SUM(load_target_nbelem))/indDLoad(*)/' "$src" > "$b"
        src="$TMP/$b"
    fi
    "$FC" -O2 -fopenmp -fPIC -c "$src" -o "${b%.*}.o"
done
"$FC" -O2 -fopenmp -fPIC -c "$HERE/ref_shim.f90" -o ref_shim.o
mkdir -p "$OUT"
"$FC" -shared -fopenmp -o "$OUT/libyetiref.so.tmp" ./*.o -Wl,--no-undefined
mv "$OUT/libyetiref.so.tmp" "$OUT/libyetiref.so"
echo "build_ref: wrote $OUT/libyetiref.so"
