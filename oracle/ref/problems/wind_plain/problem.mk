# TABULATED wiring of the line-driven-wind problem: its own Radiat() and
# CoolingSource() (Test_Problems/LineDrivenWind/cv_idl) over the generic
# ODE integrators of Src/Cooling.
CVIDL := $(REF)/Test_Problems/LineDrivenWind/cv_idl
OBJ += plm_states.o line_connect.o cooling_source.o cooling_ode_solver.o \
       radiat.o jacobian.o maxrate.o
INCLUDE_DIRS += -I$(SRC)/Cooling -I$(SRC)/LineDriven
VPATH := $(CVIDL):$(VPATH):$(SRC)/Cooling:$(SRC)/Cooling/Tabulated:$(SRC)/LineDriven
INIT_SRC := $(HERE)/common/init_wind.c
