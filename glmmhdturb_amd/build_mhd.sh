#!/usr/bin/env bash
# Build the GLM-MHD line march as a MarchBridge-compatible shared library for gfx950
# (the hipcc counterpart of march_bridge/build_mhd.sh).  The grid size N is only the
# default returned by march_nx(); any grid can be selected at run time with march_set_grid.
#
#   ./build_mhd.sh [N] [outdir]        -> <outdir>/libmhd<N>.so       (march_nv = 9)
#   NSP=2 ./build_mhd.sh [N] [outdir]  -> <outdir>/libmhd<N>_nsp2.so  (march_nv = 11: two passive
#                                         species riding the mass flux, spike_mhd.cu -DNSP=2)
#   PPM=1 ./build_mhd.sh ...             -> ..._ppm.so: the parabolic-edge build (-DPPM)
set -euo pipefail
N="${1:-480}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${2:-$HERE}"
case "${NSP:-0}" in
    0) NV=9; LIB="libmhd${N}.so" ;;
    2) NV=11; LIB="libmhd${N}_nsp2.so" ;;
    *) echo "NSP must be 0 or 2" >&2; exit 1 ;;
esac
case "${PPM:-0}" in
    0) SCHEME=0 ;;
    1) SCHEME=1; LIB="${LIB%.so}_ppm.so" ;;
    *) echo "PPM must be 0 or 1" >&2; exit 1 ;;
esac
make -s -C "$HERE" -j"${JOBS:-8}" legacy N="$N" NV="$NV" SCHEME="$SCHEME" ARCH="${ARCH:-gfx950}"
[ "$OUT" = "$HERE" ] || cp "$HERE/$LIB" "$OUT/"
echo "ok: $OUT/$LIB"
