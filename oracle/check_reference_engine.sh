#!/bin/sh
# Compile-only check of include/tr_reference_engine.hpp (the reference-side GpuTransientRenderer)
# against the reference's own headers where they lie, in the reference's default configuration
# (bunnykiller.h with Embree; Embree's headers are vendored in the reference tree). Nothing is
# linked or run. Requires the reference tree (default /root/reference); exits 0 without checking
# when it is absent.
set -e
REF=${REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
if [ ! -d "$REF/src" ]; then
  echo "check_reference_engine.sh: reference sources not found at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
cat > "$HERE/_ref/reference_engine_tu.cpp" <<'CPP'
// translation unit for the compile-only check: the reference's headers as bunnykiller.cpp includes
// them, then the adapter, instantiated for the reference's Radiance types (bunnykiller.cpp:99-117)
#include "bunnykiller.h"
#include <array>
#include <cmath>
#include <cstdio>
#include <limits>
#include <list>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#include "Utils/ProgramParams.h"
#include "Camera/PinholeCamera.h"
#include "RayTracing/World.h"
#include "Renderer/RenderEngine.h"
#include "Renderer/TransientRenderer.h"
#include "Film/StreakCameraFilm.h"
#include "Geometry/Aggregate/Mesh.h"
#include "Material/UberMaterial.h"
#include "Material/Reflectance/Lambertian.h"
#include "Material/TwoSided.h"
#include "LightSource/PointLightSource.h"
#include "LightSource/RectangularAreaLightSource.h"
#include "Sampling/StratifiedSampler.h"
#include "Integrator/BidirectionalPathTracing.h"
#include "Media/HomogeneousMedium.h"
#include "Media/PhaseFunction/IsotropicPhaseFunction.h"
#include "Utils/RandomNumbers.h"
#include "tr_reference_engine.hpp"
template class tr_reference::GpuTransientRenderer<Spectrum, Spectrum>;
template tr_scene tr_reference::make_scene<Spectrum>(const World<3, Spectrum>&, const std::vector<tr_reference::MeshEntry>&);
CPP
g++ -std=c++14 -fsyntax-only -w -I "$REF/src" -isystem "$REF/src/External" -I "$ROOT/include" \
    "$HERE/_ref/reference_engine_tu.cpp"
echo "tr_reference_engine.hpp compiles against $REF/src"
