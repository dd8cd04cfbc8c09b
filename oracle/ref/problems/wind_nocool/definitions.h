/* PLUTO compile-time configuration for the CPU oracle build: wind_nocool
 * 2-D spherical (r, theta) line-driven disk wind, PLM(van Leer) + HLLC +
 * RK2, gravity, SIROCCO-mode line force, no cooling. */
#define  PHYSICS                        HD
#define  DIMENSIONS                     2
#define  GEOMETRY                       SPHERICAL
#define  BODY_FORCE                     VECTOR
#define  COOLING                        NO
#define  RECONSTRUCTION                 LINEAR
#define  TIME_STEPPING                  RK2
#define  NTRACER                        1
#define  PARTICLES                      NO
#define  USER_DEF_PARAMETERS            10
#define  LINE_DRIVEN_WIND               SIROCCO_MODE

#define  DUST_FLUID                     NO
#define  EOS                            IDEAL
#define  ENTROPY_SWITCH                 NO
#define  INCLUDE_LES                    NO
#define  THERMAL_CONDUCTION             NO
#define  VISCOSITY                      NO
#define  RADIATION                      NO
#define  ROTATING_FRAME                 NO

#define  MU                             0
#define  T_ISO                          1
#define  L_star                         2
#define  f_x                            3
#define  f_uv                           4
#define  T_x                            5
#define  KRAD                           6
#define  ALPHARAD                       7
#define  CENT_MASS                      8
#define  GAMMA_EOS                      9

#define  WARNING_MESSAGES               NO
#define  UNIT_DENSITY                   1e-14
#define  UNIT_LENGTH                    1e9
#define  UNIT_VELOCITY                  1e9
#define  UNIT_MASS                      (UNIT_DENSITY*UNIT_LENGTH*UNIT_LENGTH*UNIT_LENGTH)
#define  UNIT_ACCELERATION              (UNIT_VELOCITY*UNIT_VELOCITY/UNIT_LENGTH)
#define  UNIT_FORCE                     (UNIT_MASS*UNIT_ACCELERATION)
#define  UNIT_TIME                      (UNIT_LENGTH/UNIT_VELOCITY)
#define  UNIT_PRESSURE                  (UNIT_DENSITY*UNIT_VELOCITY*UNIT_VELOCITY)
#define  LIMITER                        VANLEER_LIM
