#!/usr/bin/env bash
# Build the 5-variable hydro line march as a MarchBridge-compatible shared library for gfx950
# (hipcc counterpart of march_bridge/build_march.sh): exports march_nv (= 5), march_nx,
# march_set_dtdx, march_set_gamma, march_set_grid, march_run_dev and march_run_dev_f64.  N is
# the default grid (march_set_grid changes it at run time).
#
#   ./build_march.sh [N] [outdir]              -> <outdir>/libmarch<N>.so
#   SCALARS=1 ./build_march.sh [N] [outdir]    -> libmarch<N>_scalars.so (march_nv = 7, -DSCALARS)
#   SCALARS=1 CMA=1 ./build_march.sh ...       -> libmarch<N>_cma.so     (-DSCALARS -DCMA)
#   SCHEME=ppm1d | donor (with any of the above) -> ..._ppm1d.so / ..._donor.so: the parabolic-edge
#                                                (-DPPM1D) or donor-cell seam (-DDONOR) build
set -euo pipefail
N="${1:-480}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${2:-$HERE}"
case "${SCHEME:-}" in
    "") S=0; SUF="" ;;
    ppm1d) S=1; SUF="_ppm1d" ;;
    donor) S=2; SUF="_donor" ;;
    *) echo "SCHEME must be ppm1d or donor" >&2; exit 1 ;;
esac
if [ -n "${SCALARS:-}" ]; then
    C="${CMA:-0}"
    LIB="libmarch${N}_$([ "$C" = 1 ] && echo cma || echo scalars)${SUF}.so"
    make -s -C "$HERE" -j"${JOBS:-8}" legacy N="$N" NV=7 CMA="$C" SCHEME="$S" ARCH="${ARCH:-gfx950}"
else
    LIB="libmarch${N}${SUF}.so"
    make -s -C "$HERE" -j"${JOBS:-8}" legacy N="$N" NV=5 SCHEME="$S" ARCH="${ARCH:-gfx950}"
fi
[ "$OUT" = "$HERE" ] || cp "$HERE/$LIB" "$OUT/"
echo "ok: $OUT/$LIB"
