# Builds the reference MSModel sources, where they lie under /root/reference,
# into oracle/_ref/ (git-ignored): a shared library of the numerical core for
# the parity tests and the fixture generator, and the reference's own msmain
# for the CPU baseline.  Nothing is copied from the reference tree: gcc reads
# the sources in place and writes objects and binaries under oracle/_ref/.
#
# Usage (from the repository root):  make -f oracle/ref.mk
#
# The reference Makefile links -llapack, which this image lacks; the conda MKL
# (LP64, Fortran ABI) provides zgetrf_/zgetrs_/zgesv_/zgeev_/zgels_/... instead.
# -O2 is added (the reference Makefile compiles at -O0).
REF      ?= /root/reference
SRC      := $(REF)/src/core
OUT      := oracle/_ref
OBJ      := $(OUT)/obj
MKLDIR   ?= /opt/conda/lib
CC       := gcc
CFLAGS   := -O2 -fPIC -fms-extensions -fopenmp -w -I$(SRC)
LDLIBS   := -lm -L$(MKLDIR) -lmkl_rt -Wl,-rpath,$(MKLDIR)

# Numerical core on the Foldy-Lax path (medium, Green functions, MS matrix,
# incident wave, root finder) and the helpers they call.
CORE     := nru_random common_util real_vector_util complex_vector_util \
            green_bessel scattering_model_util medium_util ms_matrix_util \
            incident_wave_util domain_util root_finder
CORE_OBJ := $(patsubst %,$(OBJ)/%.o,$(CORE))
ALL_SRC  := $(wildcard $(SRC)/*.c)
ALL_OBJ  := $(patsubst $(SRC)/%.c,$(OBJ)/%.o,$(ALL_SRC))

all: $(OUT)/libmsref.so $(OUT)/libmsref_shim.so $(OUT)/libmsref_ext.so $(OUT)/msmain

# msmain_gpu needs msmodel_amd/libmsgpu.so (python msmodel_amd/build.py); not part of `all`.
gpu: $(OUT)/msmain_gpu

$(OBJ):
	mkdir -p $@

$(OBJ)/%.o: $(SRC)/%.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@

$(OBJ)/ms_main.o: $(REF)/src/main/ms_main.c | $(OBJ)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/libmsref.so: $(CORE_OBJ)
	$(CC) -shared -fopenmp $^ $(LDLIBS) -o $@

$(OUT)/libmsref_shim.so: oracle/ref_shim.c $(OUT)/libmsref.so
	$(CC) $(CFLAGS) -shared $< -L$(OUT) -lmsref -Wl,-rpath,'$$ORIGIN' $(LDLIBS) -o $@

# Every core object (grid generators, eigenstate drivers) behind the plain-C entry
# points of oracle/ref_ext_shim.c.
$(OUT)/libmsref_ext.so: oracle/ref_ext_shim.c $(ALL_OBJ)
	$(CC) $(CFLAGS) -shared $< $(ALL_OBJ) $(LDLIBS) -o $@

$(OUT)/msmain: $(ALL_OBJ) $(OBJ)/ms_main.o
	$(CC) -fopenmp $^ $(LDLIBS) -o $@

# Drop-in demonstration: the reference's own msmain with the Foldy-Lax path served by
# libmsgpu.so.  The replaced functions are made local in copies of the two reference
# objects that define them (their other functions stay), so every call from ms_core.c
# binds to the library; -lmsgpu precedes MKL so zgesv_/zgetrf_/zgetrs_ bind to it too.
# MKL is named by path: -L$(MKLDIR) would put conda's older libstdc++ in front of the
# one the HIP runtime needs.
DROPIN   := build_ms_matrix build_dk_ms_matrix build_normalized_matrix build_plane_wave \
            tracelog_lu dk_log_det_ms traceinv inverse_iteration total_cross_section \
            eval_total_wave eval_incident_wave
GPU_OBJ  := $(filter-out $(OBJ)/ms_matrix_util.o $(OBJ)/incident_wave_util.o,$(ALL_OBJ))

$(OBJ)/%.local.o: $(OBJ)/%.o
	objcopy $(foreach f,$(DROPIN),--localize-symbol=$(f)) $< $@

$(OUT)/msmain_gpu: $(GPU_OBJ) $(OBJ)/ms_matrix_util.local.o $(OBJ)/incident_wave_util.local.o $(OBJ)/ms_main.o msmodel_amd/libmsgpu.so
	$(CC) -fopenmp $(filter %.o,$^) -Lmsmodel_amd -lmsgpu -Wl,-rpath,'$$ORIGIN/../../msmodel_amd' \
	    -Wl,-rpath-link,/usr/lib/x86_64-linux-gnu -lm $(MKLDIR)/libmkl_rt.so.1 -Wl,-rpath,$(MKLDIR) -o $@

clean:
	rm -rf $(OUT)

.PHONY: all gpu clean
