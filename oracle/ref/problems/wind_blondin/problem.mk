OBJ += plm_states.o cooling.o line_connect.o
INCLUDE_DIRS += -I$(SRC)/Cooling/BLONDIN -I$(SRC)/LineDriven
VPATH := $(VPATH):$(SRC)/Cooling/BLONDIN:$(SRC)/LineDriven
INIT_SRC := $(HERE)/common/init_wind.c
