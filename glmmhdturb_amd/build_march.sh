#!/usr/bin/env bash
# Build the 5-variable hydro line march as a MarchBridge-compatible shared library for gfx950
# (hipcc counterpart of march_bridge/build_march.sh): exports march_nv (= 5), march_nx,
# march_set_dtdx, march_set_gamma, march_set_grid, march_run_dev and march_run_dev_f64.  N is
# the default grid (march_set_grid changes it at run time).
#
#   ./build_march.sh [N] [outdir]    -> <outdir>/libmarch<N>.so
set -euo pipefail
N="${1:-480}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${2:-$HERE}"
make -s -C "$HERE" -j"${JOBS:-8}" legacy N="$N" NV=5 ARCH="${ARCH:-gfx950}"
[ "$OUT" = "$HERE" ] || cp "$HERE/libmarch${N}.so" "$OUT/"
echo "ok: $OUT/libmarch${N}.so"
