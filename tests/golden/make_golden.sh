#!/bin/bash
# Regenerates the format fixtures in tests/golden/ with the reference's own
# converter (converter/converter.cpp, compiled from its sources by
# `make -C oracle ref`).  Inputs are the two text edge lists committed beside
# this script; outputs are the byte-exact .csr / .csc / .wcsr files.
set -euo pipefail
here=$(cd "$(dirname "$0")" && pwd)
make -s -C "$here/../../oracle" ref
conv="$here/../../oracle/_ref/convert"
tmp=$(mktemp -d)
trap 'rm -rf "$tmp"' EXIT
for g in tiny4 rand64; do
  cp "$here/$g.txt" "$tmp/$g.txt"
  (cd "$tmp" && "$conv" --input "$g.txt" --skip 1 --type csr >/dev/null \
             && "$conv" --input "$g.txt" --skip 1 --type csc >/dev/null \
             && "$conv" --input "$g.txt" --skip 1 --type wcsr >/dev/null)
  cp "$tmp/$g.csr" "$here/$g.csr"
  cp "$tmp/$g.csc" "$here/$g.csc"
  # the converter names this file "<base>..wcsr" (converter.cpp:777)
  cp "$tmp/$g..wcsr" "$here/$g.wcsr"
done
