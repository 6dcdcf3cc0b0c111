#!/bin/sh
# Compile the reference's statevector path from its sources into oracle/_ref/.
# Used only as a checker and as the timed CPU baseline.
set -e
cd "$(dirname "$0")"
REF="${1:-/root/reference}"
if [ ! -d "$REF/src" ]; then
    echo "reference sources not found at $REF; skipping oracle/_ref" >&2
    exit 0
fi
make -s REF="$REF" ref
