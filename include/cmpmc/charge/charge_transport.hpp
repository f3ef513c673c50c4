// Reference header path kept for source compatibility (UniformElectricField, ChargePhysics).
#pragma once
#include "cmpmc/transport/engine.hpp"
