OBJ += ppm_states.o ppm_coeffs.o
INIT_SRC := $(HERE)/common/init_file_ic.c
