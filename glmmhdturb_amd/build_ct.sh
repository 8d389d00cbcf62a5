#!/usr/bin/env bash
# Build the constrained-transport march library for gfx950 (hipcc counterpart of
# march_bridge/build_ct.sh): exports march_nv (= 8), march_nx, march_set_dtdx,
# march_set_gamma, ct_run, ct_regs, ct_shmem.  N is the default grid (march_set_grid
# changes it at run time).
#
#   ./build_ct.sh [N] [outdir]         -> <outdir>/libct<N>.so       (march_nv = 8)
#   NSP=2 ./build_ct.sh [N] [outdir]   -> <outdir>/libct<N>_nsp2.so  (march_nv = 10: two passive
#                                         species riding the mass flux, spike_ctm.cu -DNSP=2)
#   PPM=1 ./build_ct.sh ...             -> ..._ppm.so: the parabolic-edge build (-DPPM)
set -euo pipefail
N="${1:-480}"
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${2:-$HERE}"
case "${NSP:-0}" in
    0) NV=8; LIB="libct${N}.so" ;;
    2) NV=10; LIB="libct${N}_nsp2.so" ;;
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
