# TABULATED wiring of the line-driven-wind problem: its own Radiat() and
# CoolingSource() (Test_Problems/LineDrivenWind/cv_idl) over the generic
# ODE integrators of Src/Cooling.  UserDefBoundary comes from the problem's
# own init.c, compiled as it lies; its Init / InitDomain / Analysis /
# BodyForce* are made weak symbols so that those of our init_wind.c (the
# initial condition read from a file, the state dump) are the ones linked.
CVIDL := $(REF)/Test_Problems/LineDrivenWind/cv_idl
OBJ += plm_states.o line_connect.o cooling_source.o cooling_ode_solver.o \
       radiat.o jacobian.o maxrate.o
INCLUDE_DIRS += -I$(SRC)/Cooling -I$(SRC)/LineDriven
VPATH := $(CVIDL):$(VPATH):$(SRC)/Cooling:$(SRC)/Cooling/Tabulated:$(SRC)/LineDriven
INIT_SRC := $(HERE)/common/init_wind.c
OBJ += cvidl_init.o

$(OUT)/cvidl_init.o: $(CVIDL)/init.c $(PDIR)/definitions.h | $(OUT)
	$(CC) $(CFLAGS) $(INCLUDE_DIRS) $< -o $@.strong.o
	objcopy --weaken-symbol=Init --weaken-symbol=InitDomain --weaken-symbol=Analysis \
	        --weaken-symbol=BodyForceVector --weaken-symbol=BodyForcePotential $@.strong.o $@
