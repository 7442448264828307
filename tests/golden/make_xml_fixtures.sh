#!/bin/sh
# Copies the reference's example input and output files (examples/*.xmsi,
# examples/In.xmso, examples/srm1155.xmso of XMI-MSIM 8.1) into gzip'ed
# fixtures for the XMSI/XMSO tests.  Run from the repository root.
set -e
SRC=${1:-/root/reference/examples}
for f in In srm1132 srm1155 srm1412; do gzip -9 -n -c "$SRC/$f.xmsi" > tests/golden/$f.xmsi.gz; done
for f in In srm1155; do gzip -9 -n -c "$SRC/$f.xmso" > tests/golden/$f.xmso.gz; done
