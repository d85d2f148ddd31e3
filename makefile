##===----------------------------------------------------------------------===##
## Per-target fan-out over the tests (same role as the reference's makefile):
##   make gcc | make hip         build every test for one target
##   make gcc-omp                the CPU target with OpenMP (a second baseline)
##   make check.gcc | check.hip  every test once under ASan + UBSan (host code)
##   make roofline.hip           rocprofv3 traffic and bandwidth of every hip test
##   make <test>.hip             build one test
##   make test.hip               ./benchmark for the hip target
##   make benchmark              ./benchmark NX NY NS NT NRUNS hip gcc
##===----------------------------------------------------------------------===##

-include makefile.in

TARGETS = \
	divergence \
	gameoflife \
	gaussblur \
	gradient \
	jacobi \
	lapgsrb \
	laplacian \
	matmul \
	matvec \
	sincos \
	tricubic \
	tricubic2 \
	uxx1 \
	vecadd \
	wave13pt \
	whispering

.PHONY: all gcc gcc-omp hip lib clean benchmark test.gcc test.hip check.gcc check.hip roofline.hip

all: gcc hip

lib:
	$(SILENT)$(MAKE) -s -C kernelgen-perf-tests_amd ARCH=$(HIPARCH)

gcc: $(addsuffix .gcc, $(TARGETS))

gcc-omp: $(addsuffix .gcc-omp, $(TARGETS))

hip: lib $(addsuffix .hip, $(TARGETS))

%.gcc:
	$(SILENT)$(MAKE) -s -C $(subst .gcc,,$@)/gcc

%.gcc-omp:
	$(SILENT)$(MAKE) -s -C $(subst .gcc-omp,,$@)/gcc-omp

%.hip: lib
	$(SILENT)$(MAKE) -s -C $(subst .hip,,$@)/hip

check.gcc: $(addsuffix .gcc.check, $(TARGETS))

check.hip: lib $(addsuffix .hip.check, $(TARGETS))

roofline.hip: lib $(addsuffix .hip.roofline, $(TARGETS))

%.gcc.check:
	$(SILENT)$(MAKE) -s -C $(subst .gcc.check,,$@)/gcc check

%.hip.check: lib
	$(SILENT)$(MAKE) -s -C $(subst .hip.check,,$@)/hip check

%.hip.roofline: lib
	$(SILENT)$(MAKE) -s -C $(subst .hip.roofline,,$@)/hip roofline

%.gcc.clean:
	$(SILENT)$(MAKE) -s -C $(subst .gcc.clean,,$@)/gcc clean

%.gcc-omp.clean:
	$(SILENT)$(MAKE) -s -C $(subst .gcc-omp.clean,,$@)/gcc-omp clean

%.hip.clean:
	$(SILENT)$(MAKE) -s -C $(subst .hip.clean,,$@)/hip clean

clean: $(addsuffix .gcc.clean, $(TARGETS)) $(addsuffix .gcc-omp.clean, $(TARGETS)) \
	$(addsuffix .hip.clean, $(TARGETS))

test.gcc: gcc
	$(SILENT)./benchmark $(NX) $(NY) $(NS) $(NT) $(NRUNS) gcc

test.hip: hip
	$(SILENT)./benchmark $(NX) $(NY) $(NS) $(NT) $(NRUNS) hip

benchmark: hip gcc
	$(SILENT)./benchmark $(NX) $(NY) $(NS) $(NT) $(NRUNS) hip gcc
