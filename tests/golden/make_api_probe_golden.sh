#!/bin/sh
# Writes tests/golden/api_probe_{host,prepare,manager,loop,advect}.txt: the text that
# tests/native/api_probe.cpp and api_probe_advect.cpp print when built against
# the reference's own headers (oracle/_ref/awcm_api_probe_ref and
# awcm_api_probe_advect_ref, compiled in place from /root/reference by
# oracle/Makefile).  Run from anywhere in this container;
# the GPU box has no reference tree and only reads the committed text.
set -eu
here=$(cd "$(dirname "$0")" && pwd)
root=$(cd "$here/../.." && pwd)
make -s -C "$root/oracle" ref-probe
probe="$root/oracle/_ref/awcm_api_probe_ref"
tmp=$(mktemp -d)
trap 'rm -rf "$tmp"' EXIT
"$probe" host "$tmp" > "$here/api_probe_host.txt"
"$probe" prepare > "$here/api_probe_prepare.txt"
"$probe" manager > "$here/api_probe_manager.txt"
"$probe" loop 40 > "$here/api_probe_loop.txt"
"$root/oracle/_ref/awcm_api_probe_advect_ref" 40 > "$here/api_probe_advect.txt"
wc -l "$here"/api_probe_*.txt
