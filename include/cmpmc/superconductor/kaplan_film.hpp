// Reference header path kept for source compatibility (KaplanFilm configuration).
#pragma once
#include "cmpmc/geometry/surface_property.hpp"
