// planeWave (include/planeWave.hpp of the reference) lives in paraxialSource.hpp.
#pragma once
#include "paraxialSource.hpp"
