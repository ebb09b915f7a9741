#!/bin/bash
# Builds into oracle/_ref/ (test infrastructure, never committed):
#   librt_common_ref.so   the reference's host library, compiled unchanged from
#                         the reference tree in place: common-lib/beam.cpp,
#                         surface.cpp, tabulated_function.cpp, ascii_file.cpp
#                         (no third-party dependencies; scheme.cpp needs mxml and
#                         is not built; the engine symbols stay unresolved here)
#   ref_pipeline_hip      oracle/ref/ref_pipeline.cpp on that library, linked
#                         against rt_amd/librt_engine_hip.so
#   ref_pipeline_oracle   the same, linked against oracle/librt_engine_oracle.so
#   ray_tracing_hip       the reference's ray-tracing/main.cpp, unchanged, on
#                         rt_amd's Scheme (rt_amd/librt_scheme_hip.so): our
#                         include/rt/scheme.h is force-included and defines
#                         common-lib's include guards, so it takes the place of
#                         common-lib/scheme.h (which needs mxml) and its
#                         beam.h / surface.h / tabulated_function.h
#   losses_comp_hip       the reference's efficiency/losses_comp.cpp, unchanged,
#                         reading rt_amd's .dat files through its load_beam_data
#                         (include/rt/beam.h in place of common-lib/beam.h)
# Needs the reference tree (RT_REFERENCE, default /root/reference) and the two
# engine libraries already built.
set -euo pipefail
REF=${RT_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT=$ROOT/oracle/_ref
CL=$REF/common-lib
mkdir -p "$OUT"
g++ -std=c++11 -O2 -fPIC -shared -w -I"$CL" -o "$OUT/librt_common_ref.so" \
    "$CL/beam.cpp" "$CL/surface.cpp" "$CL/tabulated_function.cpp" "$CL/ascii_file.cpp"
g++ -std=c++11 -O2 -Wall -I"$CL" -o "$OUT/ref_pipeline_hip" "$HERE/ref_pipeline.cpp" \
    -L"$OUT" -lrt_common_ref -L"$ROOT/rt_amd" -lrt_engine_hip \
    -Wl,-rpath,'$ORIGIN' -Wl,-rpath,'$ORIGIN/../../rt_amd'
g++ -std=c++11 -O2 -Wall -I"$CL" -o "$OUT/ref_pipeline_oracle" "$HERE/ref_pipeline.cpp" \
    -L"$OUT" -lrt_common_ref -L"$ROOT/oracle" -lrt_engine_oracle \
    -Wl,-rpath,'$ORIGIN' -Wl,-rpath,'$ORIGIN/..'
RTLIBS="-L$ROOT/rt_amd -lrt_scheme_hip -lrt_common_hip -lrt_engine_hip -lrt_hip -Wl,-rpath,\$ORIGIN/../../rt_amd"
g++ -std=c++11 -O2 -include "$ROOT/include/rt/scheme.h" -o "$OUT/ray_tracing_hip" "$REF/ray-tracing/main.cpp" $RTLIBS
g++ -std=c++11 -O2 -include "$ROOT/include/rt/beam.h" -o "$OUT/losses_comp_hip" "$REF/efficiency/losses_comp.cpp" \
    $RTLIBS
echo "built $OUT"
