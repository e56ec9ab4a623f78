"""Physical and numerical constants (fv3core/utils/global_constants.py:1-37,
fv3core/stencils/ppm.py:8-18).  Values are data fixed by the model definition."""

RADIUS = 6.3712e6
PI = 3.1415926535897931
OMEGA = 7.2921e-5
GRAV = 9.80665
RDGAS = 287.05
RVGAS = 461.50
HLV = 2.5e6
HLF = 3.3358e5
RGRAV = 1.0 / GRAV
CP_AIR = 1004.6
KAPPA = RDGAS / CP_AIR
DZ_MIN = 2.0
CV_AIR = CP_AIR - RDGAS
RDG = -RDGAS / GRAV
CNST_0P20 = 0.2
TFREEZE = 273.15
K1K = RDGAS / CV_AIR
CV_VAP = 3.0 * RVGAS
ZVIR = RVGAS / RDGAS - 1
C_ICE = 1972.0
C_LIQ = 4.1855e3
CP_VAP = 4.0 * RVGAS
TICE = 273.16
T_MIN = 184.0
CONSV_MIN = 0.001

# PPM: volume-conserving cubic with zero 2nd derivative at the end point
PPM_C1 = -2.0 / 14.0
PPM_C2 = 11.0 / 14.0
PPM_C3 = 5.0 / 14.0
# PPM volume-mean form
PPM_P1 = 7.0 / 12.0
PPM_P2 = -1.0 / 12.0
PPM_S11 = 11.0 / 14.0
PPM_S14 = 4.0 / 7.0
PPM_S15 = 3.0 / 14.0

N_HALO_DEFAULT = 3
TRACER_VARIABLES = [
    "qvapor",
    "qliquid",
    "qrain",
    "qice",
    "qsnow",
    "qgraupel",
    "qo3mr",
    "qsgs_tke",
    "qcld",
]
