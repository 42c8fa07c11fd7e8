# Builds the plain-C oracle (test infrastructure) into oracle/libpxo.so.
#   make -f oracle/oracle.mk        (from the repository root)
CC ?= gcc
oracle/libpxo.so: oracle/pxo.c oracle/pxo.h
	$(CC) -std=c99 -O2 -fPIC -shared -ffp-contract=off -Wall -Wno-unused-function oracle/pxo.c -lm -o $@
