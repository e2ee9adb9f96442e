# Builds the reference CPU path (Fortran) from its sources where they lie,
# with AMD flang, into oracle/_ref/.  Container-only test infrastructure:
# the outputs are used to write golden fixtures and as the CPU baseline.
# Usage: make -f oracle/ref_build.mk -j8   (from the repository root)
REF  ?= /root/reference
SRC  := $(REF)/src
OUT  := oracle/_ref
FC   := flang
# The reference's serial build copies its (empty) src/config/config.h to
# src/config.h; we point the preprocessor at that same file through a symlink
# instead of writing into the reference tree.
INC  := $(OUT)/inc
FFLAGS := -O2 -ffree-form -DGNU -cpp -I$(INC)/x -I$(INC) -module-dir $(OUT)/mod -J$(OUT)/mod

MOD0 := quick_mpi_module quick_constants_module
MOD1 := quick_method_module_ quick_molspec_module_ quick_gaussian_class_module
MOD2 := quick_size_module quick_amber_interface_module quick_basis_module_ \
        quick_calculated_module_ quick_divcon_module quick_ecp_module \
        quick_electrondensity_module quick_files_module quick_gridpoints_module \
        quick_mfcc_module quick_params_module quick_pb_module quick_scratch_module \
        quick_timer_module_ quick_all_module
MODS := $(MOD0) $(MOD1) $(MOD2)
MODOBJ := $(addprefix $(OUT)/obj/m_,$(addsuffix .o,$(MODS)))

# the serial object list of the reference Makefile (Makefile:170-177, minus
# main.o, which ref_driver.f90 replaces) plus fake_amber_interface.o
CORE := initialize read_job_and_atom fmm getMolSad getMol shell schwarz \
        quick_one_electron_integral getEnergy inidivcon ecp hfoperator nuclear dft \
        sedftoperator dipole scf uscf finalize uhfoperator udftoperator usedftoperator \
        uelectdii mpi_setup quick_debug calMP2 optimize gradient hessian CPHF frequency \
        MFCC basis fake_amber_interface
COREOBJ := $(addprefix $(OUT)/obj/c_,$(addsuffix .o,$(CORE)))
SUBS := $(basename $(notdir $(wildcard $(SRC)/subs/*.f90)))
SUBOBJ := $(addprefix $(OUT)/obj/s_,$(addsuffix .o,$(SUBS)))
BLAS := $(basename $(notdir $(wildcard $(SRC)/BLAS/*.f)))
BLASOBJ := $(addprefix $(OUT)/obj/b_,$(addsuffix .o,$(BLAS)))

all: $(OUT)/qk_ref_driver

$(INC)/config.h:
	mkdir -p $(INC)/x $(OUT)/obj $(OUT)/mod
	ln -sf $(SRC)/config/config.h $(INC)/config.h

# modules must be compiled in dependency order (serially)
$(OUT)/.mods: $(INC)/config.h
	set -e; for m in $(MODS); do $(FC) $(FFLAGS) -c $(SRC)/modules/$$m.f90 -o $(OUT)/obj/m_$$m.o; done
	touch $@

$(OUT)/obj/c_%.o: $(SRC)/%.f90 $(OUT)/.mods
	$(FC) $(FFLAGS) -c $< -o $@
$(OUT)/obj/s_%.o: $(SRC)/subs/%.f90 $(OUT)/.mods
	$(FC) $(FFLAGS) -c $< -o $@
$(OUT)/obj/b_%.o: $(SRC)/BLAS/%.f $(OUT)/.mods
	$(FC) -O2 -c $< -o $@

$(OUT)/libsubs.a: $(SUBOBJ)
	rm -f $@; ar rcs $@ $^
$(OUT)/libblas.a: $(BLASOBJ)
	rm -f $@; ar rcs $@ $^

$(OUT)/qk_ref_driver: oracle/ref_driver.f90 $(COREOBJ) $(OUT)/libsubs.a $(OUT)/libblas.a
	$(FC) $(FFLAGS) oracle/ref_driver.f90 $(COREOBJ) $(OUT)/obj/m_*.o $(OUT)/libsubs.a $(OUT)/libblas.a -o $@
