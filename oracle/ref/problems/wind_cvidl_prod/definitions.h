/* PLUTO compile-time configuration for the CPU oracle build: wind_cvidl_prod
 * 2-D spherical (r, theta) line-driven disk wind, PLM(van Leer) + HLLC +
 * RK2, gravity, SIROCCO-mode line force, TABULATED cooling, and the
 * production problem's internal boundary (INTERNAL_BOUNDARY YES) and
 * userdef boundaries, taken from Test_Problems/LineDrivenWind/cv_idl/init.c.
 * The user parameters carry the cv_idl labels, in its order.  With the
 * production problem's SHOCK_FLATTENING MULTID, CHAR_LIMITING YES and
 * ENTROPY_SWITCH ALWAYS (Test_Problems/LineDrivenWind/cv_idl/definitions.h). */
#define  PHYSICS                        HD
#define  DIMENSIONS                     2
#define  GEOMETRY                       SPHERICAL
#define  BODY_FORCE                     VECTOR
#define  COOLING                        TABULATED
#define  RECONSTRUCTION                 LINEAR
#define  TIME_STEPPING                  RK2
#define  NTRACER                        1
#define  PARTICLES                      NO
#define  USER_DEF_PARAMETERS            15
#define  LINE_DRIVEN_WIND               SIROCCO_MODE

#define  DUST_FLUID                     NO
#define  EOS                            IDEAL
#define  ENTROPY_SWITCH                 ALWAYS
#define  INCLUDE_LES                    NO
#define  THERMAL_CONDUCTION             NO
#define  VISCOSITY                      NO
#define  RADIATION                      NO
#define  ROTATING_FRAME                 NO

#define  MU                             0
#define  RHO_0                          1
#define  R_0                            2
#define  RHO_ALPHA                      3
#define  CENT_MASS                      4
#define  DISK_MDOT                      5
#define  T_ISO                          6
#define  L_star                         7
#define  f_x                            8
#define  f_uv                           9
#define  T_x                            10
#define  KRAD                           11
#define  ALPHARAD                       12
#define  DFLOOR                         13
#define  GAMMA                          14

#define  WARNING_MESSAGES               NO
#define  INTERNAL_BOUNDARY              YES
#define  PMX_CVIDL_BOUNDARY             1
#define  UNIT_DENSITY                   1e-14
#define  UNIT_LENGTH                    1e9
#define  UNIT_VELOCITY                  1e9
#define  UNIT_MASS                      (UNIT_DENSITY*UNIT_LENGTH*UNIT_LENGTH*UNIT_LENGTH)
#define  UNIT_ACCELERATION              (UNIT_VELOCITY*UNIT_VELOCITY/UNIT_LENGTH)
#define  UNIT_FORCE                     (UNIT_MASS*UNIT_ACCELERATION)
#define  UNIT_TIME                      (UNIT_LENGTH/UNIT_VELOCITY)
#define  UNIT_PRESSURE                  (UNIT_DENSITY*UNIT_VELOCITY*UNIT_VELOCITY)
#define  LIMITER                        VANLEER_LIM
#define  SHOCK_FLATTENING               MULTID
#define  CHAR_LIMITING                  YES
