// Reference header path kept for source compatibility: the types live in core/core.hpp.
#pragma once
#include "cmpmc/core/core.hpp"
