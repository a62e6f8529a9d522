# Builds oracle/_ref/piny_ref: the reference's own C and Fortran sources,
# compiled where they lie under $(REF) with the reference's flags
# (compile/Linux-x86-64-gcc-serial/make_head.mk: -O2 -fPIC, include/linux),
# linked with ref_harness.c.  Test infrastructure: outputs go to _ref/ only,
# which is not committed.  The Fortran (FFTPACK and friends) is compiled with
# ROCm's flang.  Run from oracle/:  make -f ref.mk
REF    ?= /root/reference
OUT    := _ref
CC     := gcc
FC     := /opt/rocm/bin/amdflang
CFLAGS := -O2 -fPIC -DNO_CFREE -I$(REF)/include/linux -w
FFLAGS := -O2 -fPIC -w
CSRC   := $(filter-out %/main.c,$(shell find $(REF)/src -name '*.c'))
FSRC   := $(REF)/src/mathlib/fft_generic.f $(REF)/src/mathlib/math_generic.f \
          $(REF)/src/mathlib/blas_wrappers.f
COBJ   := $(addprefix $(OUT)/obj/,$(notdir $(CSRC:.c=.o)))
FOBJ   := $(addprefix $(OUT)/obj/,$(notdir $(FSRC:.f=.fo)))

vpath %.c $(sort $(dir $(CSRC)))
vpath %.f $(sort $(dir $(FSRC)))

all: $(OUT)/piny_ref $(OUT)/piny_shim

# The drop-in shim (piny_amd/shim) inside a PINY program: the reference's
# objects, the shim and its test driver, linked with the library.
SHIM   := ../piny_amd/shim
AMDLIB := ../piny_amd

$(OUT)/obj/pmx_piny.o: $(SHIM)/pmx_piny.c $(SHIM)/pmx_piny.h ../include/piny_amd.h | $(OUT)/obj
	$(CC) $(CFLAGS) -I$(SHIM) -I../include -c -o $@ $<

$(OUT)/obj/shim_driver.o: shim_driver.c $(SHIM)/pmx_piny.h | $(OUT)/obj
	$(CC) $(CFLAGS) -I$(SHIM) -I../include -c -o $@ $<

$(OUT)/piny_shim: $(COBJ) $(FOBJ) $(OUT)/obj/shim_driver.o $(OUT)/obj/pmx_piny.o $(AMDLIB)/libpiny_amd.so
	$(FC) -O2 -o $@ $(COBJ) $(FOBJ) $(OUT)/obj/shim_driver.o $(OUT)/obj/pmx_piny.o \
	    -L$(AMDLIB) -lpiny_amd -Wl,-rpath,'$$ORIGIN/../../piny_amd' -lm

$(OUT)/obj:
	mkdir -p $@

$(OUT)/obj/%.o: %.c | $(OUT)/obj
	$(CC) $(CFLAGS) -c -o $@ $<

$(OUT)/obj/%.fo: %.f | $(OUT)/obj
	$(FC) $(FFLAGS) -c -o $@ $<

$(OUT)/obj/ref_harness.o: ref_harness.c | $(OUT)/obj
	$(CC) $(CFLAGS) -c -o $@ $<

$(OUT)/piny_ref: $(COBJ) $(FOBJ) $(OUT)/obj/ref_harness.o
	$(FC) -O2 -o $@ $^ -lm

clean:
	rm -rf $(OUT)

.PHONY: all clean
