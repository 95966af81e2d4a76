# Builds the Fortran reference (venovako/JKogb) from its own sources, where they
# lie under /root/reference, into oracle/_ref/.  Nothing from the reference is
# copied into this repository: only the objects and executables land in
# oracle/_ref/ (git-ignored).  The recipe follows the reference's own makefiles
# (src/D/GNUmakefile, src/Z/GNUmakefile, src/shared/GNUmakefile) with the flag
# set of src/gnu.mk (NaN-ignoring MIN/MAX via fmin/fmax), but calls AMD flang
# directly because the stock makefiles need an external libpvn checkout.
#
#   make -f oracle/ref.mk            # all variants
#
# Variants (each in its own directory so the .mod files never mix):
#   plain : -O2 -DNDEBUG, no -march  -> FMAD(a,b,c) = a*b+c unfused (default build)
#   dbg   : -O2 without NDEBUG       -> writes DSTEP_BUILD.<step> dumps per step
#   fma   : -O2 -DNDEBUG -march=x86-64-v3 -ffp-contract=fast -> FMAD fused
REF ?= /root/reference/src
OUT ?= $(abspath $(dir $(lastword $(MAKEFILE_LIST))))/_ref
RFC ?= /opt/rocm/llvm/bin/flang
BASE = -cpp -O2 -DUSE_FLANG -DUSE_X64 -DMIND=C_FMIN -DMAXD=C_FMAX -fopenmp

SHARED = params timer utils atypes binio jstep ktransf

.PHONY: all plain dbg fma clean
all: plain dbg fma

plain: $(OUT)/plain/djk.exe $(OUT)/plain/zjk.exe $(OUT)/plain/dtest.exe $(OUT)/plain/ztest.exe
dbg:   $(OUT)/dbg/djk.exe $(OUT)/dbg/zjk.exe
fma:   $(OUT)/fma/djk.exe $(OUT)/fma/zjk.exe

FLAGS_plain = $(BASE) -DNDEBUG
FLAGS_dbg   = $(BASE)
FLAGS_fma   = $(BASE) -DNDEBUG -march=x86-64-v3 -ffp-contract=fast

# $(1) = variant
define variant
$(OUT)/$(1)/libjk.a: $(addprefix $(REF)/shared/,$(addsuffix .F90,$(SHARED)))
	mkdir -p $(OUT)/$(1)/shared
	cd $(OUT)/$(1)/shared && for m in $(SHARED); do \
	  $(RFC) $(FLAGS_$(1)) -I$(REF)/shared -c $(REF)/shared/$$$$m.F90 -o $$$$m.o || exit 1; done
	cd $(OUT)/$(1)/shared && ar rcs ../libjk.a $(addsuffix .o,$(SHARED))

$(OUT)/$(1)/djk.exe: $(OUT)/$(1)/libjk.a $(wildcard $(REF)/D/*.F90)
	mkdir -p $(OUT)/$(1)/D
	cd $(OUT)/$(1)/D && for m in dtransf dtypes dstep djk dtest; do \
	  $(RFC) $(FLAGS_$(1)) -I../shared -I$(REF)/shared -I$(REF)/D -c $(REF)/D/$$$$m.F90 -o $$$$m.o || exit 1; done
	cd $(OUT)/$(1)/D && $(RFC) $(FLAGS_$(1)) djk.o dstep.o dtransf.o dtypes.o ../libjk.a -o ../djk.exe
	cd $(OUT)/$(1)/D && $(RFC) $(FLAGS_$(1)) dtest.o dtransf.o ../libjk.a -o ../dtest.exe

$(OUT)/$(1)/dtest.exe: $(OUT)/$(1)/djk.exe

$(OUT)/$(1)/zjk.exe: $(OUT)/$(1)/libjk.a $(wildcard $(REF)/Z/*.F90)
	mkdir -p $(OUT)/$(1)/Z
	cd $(OUT)/$(1)/Z && for m in ztransf ztypes zstep zjk ztest; do \
	  $(RFC) $(FLAGS_$(1)) -I../shared -I$(REF)/shared -I$(REF)/Z -c $(REF)/Z/$$$$m.F90 -o $$$$m.o || exit 1; done
	cd $(OUT)/$(1)/Z && $(RFC) $(FLAGS_$(1)) zjk.o zstep.o ztransf.o ztypes.o ../libjk.a -o ../zjk.exe
	cd $(OUT)/$(1)/Z && $(RFC) $(FLAGS_$(1)) ztest.o ztransf.o ../libjk.a -o ../ztest.exe

$(OUT)/$(1)/ztest.exe: $(OUT)/$(1)/zjk.exe
endef

$(eval $(call variant,plain))
$(eval $(call variant,dbg))
$(eval $(call variant,fma))

clean:
	rm -rf $(OUT)
