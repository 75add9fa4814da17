// Same include path as DXMClib (dxmc_specialization.hpp:21-30, simulationpipeline.cpp:23-25);
// the facade lives in one header.
#pragma once
#include "../dxmc.hpp"
