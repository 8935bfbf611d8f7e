#!/bin/bash
# Host-only sanitizer build of the engine's host code (engine, batches, server,
# keygen, autotuner, wire formats), the C++ SPI layer and the C++ SPI suite; the gfx950
# kernel objects are linked in unchanged (GPU code is never built with a
# sanitizer).  Run from the repository root after `python -m heu_amd.build`:
#
#   bash tests/cpp/sanitize.sh address      # builds tests/cpp/build/address/
#   ASAN_OPTIONS=detect_leaks=0:protect_shadow_gap=0 tests/cpp/build/address/paillier_hip_test
#
# (needs a GPU to run).  ThreadSanitizer builds the same way ("thread"), but
# under its memory layout the HIP runtime finds no device, so it cannot run
# the suite.
set -e
kind=${1:-address}
root=$(cd "$(dirname "$0")/../.." && pwd)
out=$root/tests/cpp/build/$kind
mkdir -p "$out"
cd "$out"
CXX=/opt/rocm/llvm/bin/clang++
INC="-I$root/include -I$root/heu_amd/spi -I/opt/rocm/include -D__HIP_PLATFORM_AMD__"
FL="-O1 -g -fPIC -fno-omit-frame-pointer -pthread"
for f in engine batch server keygen tune wire; do
  hipcc --offload-arch=gfx950 -std=c++17 $FL -Xarch_host -fsanitize=$kind -c "$root/heu_amd/csrc/$f.cpp" -o $f.o
done
objs=""
for f in ops_2048 ops_3072 worklist_2048 worklist_3072 worklist_dec_2048 worklist_dec_3072 noalloc; do
  objs="$objs $root/heu_amd/_build/$f.o"
done
$CXX -shared -fPIC -o libheu_paillier.so engine.o batch.o server.o keygen.o tune.o wire.o $objs \
  -L/opt/rocm/lib -lamdhip64 -lcrypto -Wl,-rpath,/opt/rocm/lib
$CXX -std=c++20 $FL -fsanitize=$kind $INC -shared -o libheu_paillier_spi.so "$root/heu_amd/spi/paillier_hip.cc" \
  -L. -lheu_paillier -L/opt/rocm/lib -lamdhip64 -lcrypto -Wl,-rpath,'$ORIGIN' -Wl,-rpath,/opt/rocm/lib
$CXX -std=c++20 $FL -fsanitize=$kind $INC -o paillier_hip_test "$root/tests/cpp/paillier_hip_test.cc" \
  -L. -lheu_paillier_spi -lheu_paillier -L/opt/rocm/lib -lamdhip64 -lcrypto -Wl,-rpath,'$ORIGIN' \
  -Wl,-rpath,/opt/rocm/lib
rm -f ./*.o
echo "built $out/paillier_hip_test"
