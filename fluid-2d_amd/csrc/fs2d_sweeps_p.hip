// Pressure-only Gauss-Seidel unit launches (launch_pressure_units): the
// sweep kernels' source compiled a second time with only that dispatch, so the
// Makefile can give these kernels their own machine-scheduler choice.
#define FS2D_PRESSURE_UNITS_ONLY 1
#include "fs2d_sweeps.hip"
