// CurvedWaveGuideFD (include/curvedWaveGuide2D.hpp of the reference) lives in waveGuideFDSimulation.hpp.
#pragma once
#include "waveGuideFDSimulation.hpp"
