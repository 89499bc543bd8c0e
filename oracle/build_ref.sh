#!/bin/sh
# Compiles the reference's own host-side functions on the SpMV path into
# oracle/_ref/libref_host.so, straight from the reference tree (nothing is
# copied into this repository).  The reference files as a whole need the CUDA
# runtime and cuSPARSE headers, which this image lacks; the functions below are
# plain C++ and are compiled from their line ranges, piped to g++:
#   get_row_from_index   src/spmv_helper.cu:18-44
#   findFirstInSorted    src/spmv_helper.cu:646-664
#   findLastInSorted     src/spmv_helper.cu:667-685
#   csr_spmv_cpu         test/dspmv_test_smt.cu:39-52
#   coo_spmv_cpu         test/dspmv_test_dgx1.cu:55-69
# Each range is checked to start with the expected function before use.
set -eu
REF=${SBLAS_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
[ -d "$REF" ] || { echo "reference tree not present: $REF" >&2; exit 3; }
mkdir -p "$OUT"

span() { # file first last expected-name
  line=$(sed -n "$2p" "$REF/$1")
  case "$line" in
    *"$4"*) ;;
    *) echo "build_ref: $1:$2 does not start $4 (got: $line)" >&2; exit 4 ;;
  esac
  sed -n "$2,$3p" "$REF/$1"
}

{
  echo '#include <cmath>'
  span src/spmv_helper.cu 18 44 get_row_from_index
  span src/spmv_helper.cu 646 664 findFirstInSorted
  span src/spmv_helper.cu 667 685 findLastInSorted
  span test/dspmv_test_smt.cu 39 52 csr_spmv_cpu
  span test/dspmv_test_dgx1.cu 55 69 coo_spmv_cpu
} | g++ -x c++ -O2 -fPIC -shared -ffp-contract=off -w -o "$OUT/libref_host.so" -
echo "built $OUT/libref_host.so"
