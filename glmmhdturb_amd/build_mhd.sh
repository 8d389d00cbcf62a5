#!/usr/bin/env bash
# Build the GLM-MHD line march as a MarchBridge-compatible shared library for gfx950
# (the hipcc counterpart of march_bridge/build_mhd.sh).  The grid size N is only the
# default returned by march_nx(); any grid can be selected at run time with march_set_grid.
#
#   ./build_mhd.sh [N] [outdir]      -> <outdir>/libmhd<N>.so
set -euo pipefail
N="${1:-480}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${2:-$HERE}"
make -s -C "$HERE" -j"${JOBS:-8}" legacy N="$N" NV=9 ARCH="${ARCH:-gfx950}"
[ "$OUT" = "$HERE" ] || cp "$HERE/libmhd${N}.so" "$OUT/"
echo "ok: $OUT/libmhd${N}.so"
