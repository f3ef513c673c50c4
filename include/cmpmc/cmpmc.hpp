// Umbrella header of the drop-in cmpmc C++ API (the reference's cmpmc/cmpmc.hpp names
// for the track-transport path), backed by libcmpmc_amd (include/cmpmc_amd.h).
#pragma once

#include "cmpmc/core/core.hpp"
#include "cmpmc/materials/crystal_config.hpp"
#include "cmpmc/lattice/crystal_lattice.hpp"
#include "cmpmc/geometry/electrode_pattern.hpp"
#include "cmpmc/geometry/mesh_field.hpp"
#include "cmpmc/phonon/interface_transmission.hpp"
#include "cmpmc/geometry/surface_property.hpp"
#include "cmpmc/superconductor/kaplan_film.hpp"
#include "cmpmc/transport/engine.hpp"
#include "cmpmc/transport/parallel.hpp"
#include "cmpmc/charge/niel_models.hpp"
#include "cmpmc/charge/energy_partition.hpp"
