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
	$(SILENT)rm -f $(TEST) *.o

ifeq ($(DIMS), 2)
test: $(TEST)
	$(SILENT)./$< $(NX) $$(($(NY) * $(NS))) $(NT)
else
test: $(TEST)
	$(SILENT)./$< $(NX) $(NY) $(NS) $(NT)
endif

.PHONY: all clean test
