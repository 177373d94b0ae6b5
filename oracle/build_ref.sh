#!/bin/sh
# Builds oracle/_ref/json_dump: oracle/json_dump.cpp against the reference's vendored JSON library
# header (third_party/nlohmann_json/include), read where it lies. Test infrastructure only; where
# the reference tree is absent there is nothing to build.
set -e
here=$(cd "$(dirname "$0")" && pwd)
ref=${AUTAN_REFERENCE:-/root/reference}
inc="$ref/third_party/nlohmann_json/include"
if [ ! -f "$inc/nlohmann/json.hpp" ]; then
  echo "build_ref.sh: no $inc/nlohmann/json.hpp, oracle/_ref not built"
  exit 0
fi
mkdir -p "$here/_ref"
g++ -std=c++17 -O2 -I "$inc" -o "$here/_ref/json_dump" "$here/json_dump.cpp"
