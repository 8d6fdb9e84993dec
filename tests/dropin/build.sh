#!/usr/bin/env bash
# Build tests/dropin/dropin_check (the Fortran caller the GPU drop-in tests
# run) from this repository's sources only, against the in-tree
# libmixsolver_hip.so.  Objects and module files go to a temporary directory.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
FDIR="$(cd "$HERE/../../mixsolver_amd/fortran" && pwd)"
LIBDIR="$(cd "$HERE/../../mixsolver_amd" && pwd)"
FC=${FC:-amdflang}
if ! command -v "$FC" >/dev/null 2>&1; then
  echo "no Fortran compiler ($FC); dropin_check not built" >&2
  exit 0
fi
TMP="$(mktemp -d /tmp/mixsolver_dropin.XXXXXX)"
trap 'rm -rf "$TMP"' EXIT
cd "$TMP"
for s in "$HERE/mod_precision.f90" "$FDIR/mixsolver_c.f90" "$FDIR/mixsolver_dropin.f90" \
         "$FDIR/mod_uns_linsolver.f90" "$FDIR/mod_uns_linsolver_mpi.f90" "$HERE/dropin_check.f90"; do
  "$FC" -O2 -c "$s"
done
"$FC" -O2 -o "$HERE/dropin_check.tmp" mod_precision.o mixsolver_c.o mixsolver_dropin.o mod_uns_linsolver.o \
  mod_uns_linsolver_mpi.o dropin_check.o -L"$LIBDIR" -lmixsolver_hip -Wl,-rpath,'$ORIGIN/../../mixsolver_amd'
mv "$HERE/dropin_check.tmp" "$HERE/dropin_check"
