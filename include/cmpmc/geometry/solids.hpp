// Reference header path kept for source compatibility (CylinderSolid, BoundaryHit).
#pragma once
#include "cmpmc/geometry/surface_property.hpp"
