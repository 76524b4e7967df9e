#!/usr/bin/env bash
# build_ref.sh -- compile the reference program's serial path, configured.
#
# TEST INFRASTRUCTURE ONLY: the binary it makes checks the CPU oracle.
#
# The reference fixes its grid size, boundary condition, loop counts and MPI
# use with unconditional #defines and const locals (phasefield.c:13-15,
# 47-51, 104), so -D flags cannot override them.  This recipe streams
# /root/reference/phasefield.c through sed into gcc's stdin: the source is
# compiled where it lies and no copy of it is written anywhere.  Only the
# binary lands, in oracle/_ref/.
#
#   build_ref.sh NX NY NZ BC NITER NSAVE ETAL OUT
#     BC    0 periodic, 1 no-flux          (phasefield.c:15)
#     ETAL  0 as shipped (etaL never read, phasefield.c:1032)
#           1 the /etaL read goes through the /etaL id
#     OUT   output binary name under oracle/_ref/
#   REF_NITER_INIT (env, optional) replaces n_iter_init (phasefield.c:105),
#     the step from which only c is swept (phasefield.c:214-218).
#
# Needs gcc and a serial HDF5 (headers and libhdf5).  HDF5_DIR defaults to
# /opt/conda.
set -euo pipefail
if [ $# -ne 8 ]; then
    echo "usage: $0 NX NY NZ BC NITER NSAVE ETAL OUT" >&2
    exit 2
fi
NX=$1 NY=$2 NZ=$3 BC=$4 NITER=$5 NSAVE=$6 ETAL=$7 OUT=$8
REF=${REF_SRC:-/root/reference/phasefield.c}
HDF5_DIR=${HDF5_DIR:-/opt/conda}
HERE=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$HERE/_ref"
[ -r "$REF" ] || { echo "reference source $REF not readable" >&2; exit 3; }

SED=(-e 's/^#define USE_MPI 1$/#define USE_MPI 0/'
     -e "s/^#define USE_BC_TYPE 1 /#define USE_BC_TYPE ${BC} /"
     -e "s/^#define NX 100$/#define NX ${NX}/"
     -e "s/^#define NY 100$/#define NY ${NY}/"
     -e "s/^#define NZ 100$/#define NZ ${NZ}/"
     -e "s/const long n_iter=5000001L,n_save=50000L;/const long n_iter=${NITER}L,n_save=${NSAVE}L;/")
if [ -n "${REF_NITER_INIT:-}" ]; then
    SED+=(-e "s/const long n_iter_init = 100000000L;/const long n_iter_init = ${REF_NITER_INIT}L;/")
fi
if [ "$ETAL" = 1 ]; then
    SED+=(-e 's/H5Dread(dataset_id_S,H5T_IEEE_F64LE,memspace_id,filespace_id,H5P_DEFAULT,etaL)/H5Dread(dataset_id_L,H5T_IEEE_F64LE,memspace_id,filespace_id,H5P_DEFAULT,etaL)/')
fi

# Check every edit applied exactly as many times as expected.
CFG=$(sed "${SED[@]}" "$REF")
check() { local n; n=$(grep -c -- "$1" <<<"$CFG" || true); [ "$n" = "$2" ] || { echo "config edit '$1' matched $n times, expected $2" >&2; exit 4; }; }
check '^#define USE_MPI 0$' 1
check "^#define USE_BC_TYPE ${BC} " 1
check "^#define NX ${NX}$" 1
check "n_iter=${NITER}L,n_save=${NSAVE}L" 1
if [ -n "${REF_NITER_INIT:-}" ]; then check "n_iter_init = ${REF_NITER_INIT}L;" 1; fi
if [ "$ETAL" = 1 ]; then check 'H5Dread(dataset_id_L,' 1; fi

# -O2 without -march: x86-64 baseline has no FMA, so nothing is contracted.
# -mcmodel=medium lets the nine static arrays exceed 2 GiB at large sizes.
sed "${SED[@]}" "$REF" | gcc -O2 -w -mcmodel=medium -x c - -o "$HERE/_ref/$OUT" \
    -I"$HDF5_DIR/include" -L"$HDF5_DIR/lib" -Wl,-rpath,"$HDF5_DIR/lib" -lhdf5 -lm
echo "$HERE/_ref/$OUT"
