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
	$(SILENT)rm -f $(TEST) $(TEST)_check *.o

ifeq ($(DIMS), 2)
RUNARGS = $(NX) $$(($(NY) * $(NS))) $(NT)
else
RUNARGS = $(NX) $(NY) $(NS) $(NT)
endif

test: $(TEST)
	$(SILENT)./$< $(RUNARGS)

# check: the program with the CPU backend, every C file built with
# AddressSanitizer and UndefinedBehaviorSanitizer, run once (the reference's
# cuda-memcheck rule has no CPU counterpart; this is the host-side one).
SANFLAGS = -g -O1 -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=undefined
$(TEST)_check: ../$(TEST).c $(HOSTDIR)/kg_host.c $(HOSTDIR)/kg_host.h $(HOSTDIR)/kg_cpu.c \
		$(HOSTDIR)/kg_cpu_whispering.c ../../include/kg_stencil.h
	$(SILENT)gcc -D_GNU_SOURCE -std=c99 $(SANFLAGS) $(OMPFLAGS) ../$(TEST).c $(HOSTDIR)/kg_host.c \
		$(HOSTDIR)/kg_cpu.c $(HOSTDIR)/kg_cpu_whispering.c -o $@ -lrt -lm

check: $(TEST)_check
	$(SILENT)ASAN_OPTIONS=detect_leaks=1 UBSAN_OPTIONS=print_stacktrace=1 ./$< $(RUNARGS)

.PHONY: all clean test check
