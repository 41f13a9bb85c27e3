#!/bin/bash
# Builds the drop-in example against include/ and rti_amd/librti_hip.so.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
g++ -std=c++17 -O2 -Wall -I"$ROOT/include" "$HERE/dropin_trench.cpp" -o "$HERE/dropin_trench" \
  -L"$ROOT/rti_amd" -lrti_hip -Wl,-rpath,"$ROOT/rti_amd"
g++ -std=c++17 -O2 -Wall "$HERE/vtp_probe.cpp" -o "$HERE/vtp_probe" -lz
g++ -std=c++17 -O2 -Wall -I"$ROOT/include" "$HERE/dropin_plugins.cpp" -o "$HERE/dropin_plugins" \
  -L"$ROOT/rti_amd" -lrti_hip -Wl,-rpath,"$ROOT/rti_amd"
