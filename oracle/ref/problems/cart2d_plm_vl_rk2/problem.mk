OBJ += plm_states.o
INIT_SRC := $(HERE)/common/init_file_ic.c
