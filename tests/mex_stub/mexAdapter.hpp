// Test stand-in for MATLAB's mexAdapter.hpp (entry-point glue); syntax checks only.
#pragma once
#include "mex.hpp"
