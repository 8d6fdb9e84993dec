#!/usr/bin/env bash
# Build the reference's own serial FP64 Fortran path (test infrastructure).
#
# Compiles the reference modules where they lie under /root/reference with
# amdflang -O2 (no -march: baseline x86-64, so no FMA contraction), links them
# with oracle/ref_harness.f90 and writes ONE executable: oracle/_ref/ref_harness.
# Intermediate objects/modules live in a temporary directory outside the
# repository and are deleted afterwards.
#
# One build-time fix is applied to a temporary copy of
# src/common/mod_vtu_writer.f90 (an output-only module pulled in through
# mod_uns_output): flang rejects the second PUBLIC on `vtu_data` (line 78)
# that gfortran tolerates; the token `, public` is dropped there.  Nothing
# else is changed, stubbed or substituted.
set -euo pipefail
REF=${MIXSOLVER_REFERENCE:-/root/reference}/src
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
FC=${FC:-amdflang}
if [ ! -d "$REF/unstructured" ]; then
  echo "reference sources not present; skipping reference build" >&2
  exit 0
fi
mkdir -p "$OUT"
TMP="$(mktemp -d /tmp/mixsolver_refbuild.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT
sed 's/type, public :: vtu_data/type :: vtu_data/' \
    "$REF/common/mod_vtu_writer.f90" > "$TMP/mod_vtu_writer.f90"
SRCS=(
  "$REF/common/mod_precision.f90"
  "$REF/common/mod_constants.f90"
  "$TMP/mod_vtu_writer.f90"
  "$REF/common/mod_cell_types.f90"
  "$REF/common/mod_interface.f90"
  "$REF/unstructured/mod_uns_mesh.f90"
  "$REF/unstructured/mod_uns_cas_reader.f90"
  "$REF/unstructured/mod_uns_connectivity.f90"
  "$REF/unstructured/mod_uns_control.f90"
  "$REF/unstructured/mod_uns_geometry.f90"
  "$REF/unstructured/mod_uns_bc.f90"
  "$REF/unstructured/mod_uns_fields.f90"
  "$REF/unstructured/mod_uns_linsolver.f90"
  "$REF/unstructured/mod_uns_output.f90"
  "$REF/unstructured/mod_uns_simple.f90"
  "$REF/unstructured/mod_uns_forces.f90"
)
OBJS=()
cd "$TMP"
for s in "${SRCS[@]}"; do
  o="$TMP/$(basename "${s%.f90}").o"
  "$FC" -cpp -O2 -c "$s" -o "$o"
  OBJS+=("$o")
done
"$FC" -cpp -O2 -c "$HERE/ref_harness.f90" -o "$TMP/ref_harness.o"
"$FC" -O2 -o "$OUT/ref_harness" "$TMP/ref_harness.o" "${OBJS[@]}"
echo "built $OUT/ref_harness"

# Drop-in compile-and-link check: the same chain with the reference's
# mod_uns_linsolver.f90 replaced by mixsolver_amd/fortran/mod_uns_linsolver.f90
# (whose solvers call libmixsolver_hip.so), plus the one-partition
# mod_uns_linsolver_mpi.f90 drop-in compiled against the reference's
# mod_precision.  oracle/_ref/ref_dropin links the reference's own drivers to
# the device solvers.  It is a reference-derived binary: it stays on the
# machine that built it (oracle/_ref is never shipped with the GPU runs) and
# no GPU test runs it; tests/dropin/dropin_check is the caller the GPU tests
# run.
LIBDIR="$(cd "$HERE/../mixsolver_amd" && pwd)"
if [ -f "$LIBDIR/libmixsolver_hip.so" ]; then
  D="$TMP/dropin"
  mkdir -p "$D"
  cd "$D"
  "$FC" -cpp -O2 -c "$LIBDIR/fortran/mixsolver_c.f90" -o "$D/mixsolver_c.o"
  "$FC" -cpp -O2 -c "$LIBDIR/fortran/mixsolver_dropin.f90" -o "$D/mixsolver_dropin.o"
  DOBJS=("$D/mixsolver_c.o" "$D/mixsolver_dropin.o")
  for s in "${SRCS[@]}"; do
    [ "$(basename "$s")" = mod_uns_linsolver.f90 ] && s="$LIBDIR/fortran/mod_uns_linsolver.f90"
    o="$D/$(basename "${s%.f90}").o"
    "$FC" -cpp -O2 -c "$s" -o "$o"
    DOBJS+=("$o")
  done
  "$FC" -cpp -O2 -c "$LIBDIR/fortran/mod_uns_linsolver_mpi.f90" -o "$D/mod_uns_linsolver_mpi.o"
  DOBJS+=("$D/mod_uns_linsolver_mpi.o")
  # the loop-level drop-in: the case adapter, compiled against the
  # reference's modules, and the partitioned-driver glue, type-checked
  # against the local-mesh components it reads (the reference's own
  # mod_uns_local_mesh needs MPI); case_dump writes what the adapter hands
  # to the library, for tests/test_oracle_golden.py
  "$FC" -cpp -O2 -c "$LIBDIR/fortran/mixsolver_case.f90" -o "$D/mixsolver_case.o"
  DOBJS+=("$D/mixsolver_case.o")
  mkdir -p "$D/glue"
  ( cd "$D/glue" && "$FC" -cpp -O2 -I"$D" -c "$HERE/../tests/dropin/local_mesh_iface.f90" \
      && "$FC" -cpp -O2 -I"$D" -c "$LIBDIR/fortran/mod_uns_simple_mpi.f90" )
  "$FC" -cpp -O2 -c "$HERE/../tests/dropin/case_dump.f90" -o "$D/case_dump.o"
  "$FC" -O2 -o "$OUT/case_dump" "$D/case_dump.o" "${DOBJS[@]}" -L"$LIBDIR" -lmixsolver_hip \
    -Wl,-rpath,'$ORIGIN/../../mixsolver_amd'
  echo "built $OUT/case_dump"
  "$FC" -cpp -O2 -c "$HERE/ref_harness.f90" -o "$D/ref_harness.o"
  "$FC" -O2 -o "$OUT/ref_dropin" "$D/ref_harness.o" "${DOBJS[@]}" -L"$LIBDIR" -lmixsolver_hip \
    -Wl,-rpath,'$ORIGIN/../../mixsolver_amd'
  echo "built $OUT/ref_dropin"
fi
