# Rules of a <test>/hip target directory.  The including makefile sets TEST
# (program name) and DIMS (3: argv nx ny ns nt; 2: argv nx ny*ns nt, as the
# reference's 2D tests are run).
-include ../../makefile.in

PKG = ../../kernelgen-perf-tests_amd
HOSTDIR = $(PKG)/host
LIBDIR = $(PKG)/lib
HOSTSRC = $(HOSTDIR)/kg_host.c $(HOSTDIR)/kg_target_hip.c

all: $(TEST)

$(LIBDIR)/libkgstencil.so:
	$(SILENT)$(MAKE) -s -C $(PKG) ARCH=$(HIPARCH)

$(TEST): ../$(TEST).c $(HOSTSRC) $(HOSTDIR)/kg_host.h ../../include/kg_stencil.h $(LIBDIR)/libkgstencil.so
	$(SILENT)$(HOSTCC) ../$(TEST).c $(HOSTSRC) -o $@ -L$(LIBDIR) -lkgstencil \
		-Wl,-rpath,'$$ORIGIN/../../kernelgen-perf-tests_amd/lib' -lrt

clean:
	$(SILENT)rm -f $(TEST) $(TEST)_check *.o

ifeq ($(DIMS), 2)
RUNARGS = $(NX) $$(($(NY) * $(NS))) $(NT)
else
RUNARGS = $(NX) $(NY) $(NS) $(NT)
endif

test: $(TEST)
	$(SILENT)./$< $(RUNARGS)

# check: the host program and the host helpers built with AddressSanitizer
# and UndefinedBehaviorSanitizer and run once -- host code only (there is no
# GPU sanitizer here), in the role of the reference's cuda-memcheck rule
# (wave13pt/cuda/makefile:54-55).
SANFLAGS = -g -O1 -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=undefined
$(TEST)_check: ../$(TEST).c $(HOSTSRC) $(HOSTDIR)/kg_host.h ../../include/kg_stencil.h $(LIBDIR)/libkgstencil.so
	$(SILENT)$(HOSTCC) $(SANFLAGS) ../$(TEST).c $(HOSTSRC) -o $@ -L$(LIBDIR) -lkgstencil \
		-Wl,-rpath,'$$ORIGIN/../../kernelgen-perf-tests_amd/lib' -lrt

check: $(TEST)_check
	$(SILENT)ASAN_OPTIONS=detect_leaks=0:verify_asan_link_order=0 UBSAN_OPTIONS=print_stacktrace=1 \
		./$< $(RUNARGS)

# roofline: per-launch kernel times (as the test prints them), then
# rocprofv3's kernel trace and one FETCH_SIZE and one WRITE_SIZE pass, summed
# up per kernel by kgprof.py: mean duration, algorithmic bytes and GB/s
# against 8 TB/s, and the measured HBM bytes per launch -- the reference's
# roofline rule reports DRAM throughput the same way
# (wave13pt/cuda/makefile:57-65).
PROFDIR ?= /tmp/kg_roofline_$(TEST)
ROCPROF ?= rocprofv3
roofline: $(TEST)
	$(SILENT)PROFILING_FNAME=`cat kernel` ./$< $(RUNARGS)
	$(SILENT)rm -rf $(PROFDIR)
	$(SILENT)cd /tmp && TMPDIR=/tmp $(ROCPROF) --kernel-trace --output-format csv -d $(PROFDIR) \
		-o trace -- $(CURDIR)/$< $(RUNARGS) > /dev/null
	$(SILENT)cd /tmp && TMPDIR=/tmp $(ROCPROF) --pmc FETCH_SIZE --output-format csv -d $(PROFDIR) \
		-o fetch -- $(CURDIR)/$< $(RUNARGS) > /dev/null
	$(SILENT)cd /tmp && TMPDIR=/tmp $(ROCPROF) --pmc WRITE_SIZE --output-format csv -d $(PROFDIR) \
		-o write -- $(CURDIR)/$< $(RUNARGS) > /dev/null
	$(SILENT)python3 $(PKG)/kgprof.py --trace $(PROFDIR)/trace_kernel_trace.csv \
		--fetch $(PROFDIR)/fetch_counter_collection.csv --write $(PROFDIR)/write_counter_collection.csv \
		--scale 2.0 1.0 --nx $$(($(NX) * $(NY) * $(NS))) --cmd "make roofline NX=$(NX) NY=$(NY) NS=$(NS) NT=$(NT)"

.PHONY: all clean test check roofline
