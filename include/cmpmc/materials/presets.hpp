// Reference header path kept for source compatibility (materials::Germanium, AluminumFilm).
#pragma once
#include "cmpmc/materials/crystal_config.hpp"
