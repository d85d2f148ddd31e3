# Rules of a <test>/gcc target directory: the same host program linked with
# the CPU backend (kg_cpu.c), built with the reference's gcc flags.  The
# including makefile sets TEST and DIMS (see target_hip.mk).  <test>/gcc-omp
# sets OMPFLAGS = -fopenmp: the backend's loops carry the reference's
# `#pragma omp parallel for` (e.g. laplacian/laplacian.c:78-80), so that
# build runs them on every host core (whispering's sweep stays serial).
-include ../../makefile.in
OMPFLAGS ?=

HOSTDIR = ../../kernelgen-perf-tests_amd/host
# compiler of the host program (whispering overrides it with GCC_O0)
MAINCC ?= $(GCC)

all: $(TEST)

kg_cpu.o: $(HOSTDIR)/kg_cpu.c ../../include/kg_stencil.h
	$(SILENT)$(GCC) $(OMPFLAGS) -c $< -o $@

kg_cpu_whispering.o: $(HOSTDIR)/kg_cpu_whispering.c ../../include/kg_stencil.h
	$(SILENT)$(GCC_O0) -c $< -o $@

$(TEST): ../$(TEST).c $(HOSTDIR)/kg_host.c $(HOSTDIR)/kg_host.h kg_cpu.o kg_cpu_whispering.o
	$(SILENT)$(MAINCC) $(OMPFLAGS) ../$(TEST).c $(HOSTDIR)/kg_host.c kg_cpu.o kg_cpu_whispering.o -o $@ -lrt -lm

clean:
	$(SILENT)rm -f $(TEST) *.o

ifeq ($(DIMS), 2)
test: $(TEST)
	$(SILENT)./$< $(NX) $$(($(NY) * $(NS))) $(NT)
else
test: $(TEST)
	$(SILENT)./$< $(NX) $(NY) $(NS) $(NT)
endif

.PHONY: all clean test
