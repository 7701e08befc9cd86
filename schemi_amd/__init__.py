"""MI355X-native explicit finite-volume step of the Schemi solver."""
from .enums import (boundaryConditionType, flowSolver, typeOfTVDLimiter, errors,
                    gasModel, typeOfSolverEnum, dimensions, timestep, turbulenceModel)
from .case import Case, BC, Turbulence
