# The line-driven wind without cooling (COOLING NO): the line force and the
# hydro step only.
OBJ += plm_states.o line_connect.o
INCLUDE_DIRS += -I$(SRC)/LineDriven
VPATH := $(VPATH):$(SRC)/LineDriven
INIT_SRC := $(HERE)/common/init_wind.c
