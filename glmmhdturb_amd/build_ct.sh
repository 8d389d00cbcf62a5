#!/usr/bin/env bash
# Build the constrained-transport march library for gfx950 (hipcc counterpart of
# march_bridge/build_ct.sh): exports march_nv (= 8), march_nx, march_set_dtdx,
# march_set_gamma, ct_run, ct_regs, ct_shmem.  N is the default grid (march_set_grid
# changes it at run time).
#
#   ./build_ct.sh [N] [outdir]       -> <outdir>/libct<N>.so
set -euo pipefail
N="${1:-480}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${2:-$HERE}"
make -s -C "$HERE" -j"${JOBS:-8}" legacy N="$N" NV=8 ARCH="${ARCH:-gfx950}"
[ "$OUT" = "$HERE" ] || cp "$HERE/libct${N}.so" "$OUT/"
echo "ok: $OUT/libct${N}.so"
