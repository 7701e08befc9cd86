"""Enumerations of the reference solver's step-solver API.

Names and integer values are those of the reference's ``enum class`` types, so
that the C-ABI, the C++ host adapter and Python all exchange the same codes:

* ``boundaryConditionType``  -- Schemi/enumerations/boundaryConditionTypesEnum.hpp:15-34
* ``flowSolver``             -- Schemi/enumerations/flowSolverEnum.hpp:15-18
* ``typeOfTVDLimiter``       -- Schemi/enumerations/typeOfTVDLimiterEnum.hpp:15-32
* ``errors``                 -- Schemi/enumerations/errorsEnum.hpp:15-30
* ``gasModel``               -- Schemi/enumerations/gasModelEnum.hpp:15-18
* ``typeOfSolverEnum``       -- Schemi/enumerations/typeOfSolverEnum.hpp:13-16
* ``dimensions``             -- Schemi/enumerations/dimensionsEnum.hpp:15-18
* ``timestep``               -- Schemi/enumerations/timestepEnum.hpp:13-18

``boundaryConditionType.periodic`` has no counterpart in the reference (which
has no periodic boundary); its value is placed after the reference's last one.
"""
from enum import IntEnum


class boundaryConditionType(IntEnum):
    blank = 0
    freeBoundary = 1
    slip = 2
    fixedValueCell = 3
    fixedValueSurface = 4
    innerSurface = 5
    calculated = 6
    calculatedTurbulentViscosity = 7
    calculatedTemperature = 8
    calculatedMassFraction = 9
    calculatedAverageMolarMass = 10
    calculatedNonidealityCorrectionPerDensity = 11
    calculatedCv = 12
    calculatedCvM = 13
    calculatedDensity = 14
    calculatedParallelBoundary = 15
    calculatedMolarFraction = 16
    periodic = 32


class flowSolver(IntEnum):
    HLL = 0
    HLLCF = 1
    HLLC = 2
    HLLCLM = 3
    KT = 4
    HLLC2p = 5
    Richtmyer = 6


class typeOfTVDLimiter(IntEnum):
    zero = 0
    minmod = 1
    vanLeer = 2
    linear = 3
    vanAlbada = 4
    HQUICK = 5
    vanLeer2 = 6
    superbee = 7
    vanAlbada2 = 8
    minmod2 = 9
    Sweby = 10
    Koren = 11
    Koren2 = 12
    UTCDFS = 13
    Schmidtmann = 14


class errors(IntEnum):
    noErrors = 0
    RiemannSolverError = 1
    negativeTemperatureError = 2
    tooBigOutputNumberError = 3
    systemError = 4
    boundaryConditionError = 5
    cubicEquationError = 6
    positivnessError = 7
    MPIError = 8
    fieldInitialisationError = 9
    meshGenerationError = 10
    initialisationError = 11
    NaNError = 12


class gasModel(IntEnum):
    ideal = 0
    vanDerWaals = 1
    RedlichKwong = 2
    stiffened = 3
    KataokaVanDerWaals = 4


class typeOfSolverEnum(IntEnum):
    SecondOrder = 0
    SecondOrderRK = 1
    ThirdOrder = 2
    ThirdOrderCada = 3


class dimensions(IntEnum):
    task1D = 0
    task2D = 1
    task3D = 2


class turbulenceModel(IntEnum):
    zeroModel = 0
    decayModel = 1
    shearModel = 2
    arithmeticA1Model = 3
    arithmeticA2Model = 4
    arithmeticA3Model = 5
    kEpsAModel = 6
    BHRModel = 7
    BHRKLModel = 8
    unknownModel = 9


class timestep(IntEnum):
    CourantTimeStep = 0
    CourantAndSourceTimeStep = 1
    CourantAndSourceAndDiffusionTimeStep = 2


# Patch order of the reference's boundary files and face blocks:
# tail (x = 0), point (x = L), bottom (z = 0), right (y = 0), left (y = L),
# top (z = L)  (Schemi/main/Schemi.cpp:575-592).
PATCHES = ("tail", "point", "bottom", "right", "left", "top")
