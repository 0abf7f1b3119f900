# Recipe that compiles the reference GiftBTE sources (read in place under
# REF, never copied) plus this repo's own dump driver into oracle/_ref/.
#
#   make -f oracle/ref.mk            # from the repo root
#
# Outputs (all under oracle/_ref/, git-ignored):
#   obj/*.o        reference translation units (src/CMakeLists.txt list,
#                  minus BTEMain.cpp and TransientFourier, which is not built
#                  upstream either: src/CMakeLists.txt:19)
#   BTE_CPU        the reference's own steady/transient driver (src/BTEMain.cpp)
#   ref_dump       oracle/ref_dump.cpp linked against the reference objects:
#                  runs the steady StaticBTESynthetic path at np=1 and writes
#                  every field at full precision (.npy)
#   mpilib/        symlinks to the MPICH runtime of this image (the conda
#                  mpicxx wrapper is broken, and rpath-ing /opt/conda/lib itself
#                  pulls in an older libstdc++)
#
# The reference's CMakeLists declares Fortran (CMakeLists.txt:2) and cannot be
# configured here; this recipe compiles the same file list directly with g++.

REF      ?= /root/reference
MPIROOT  ?= /opt/conda
OUT      := oracle/_ref
OBJ      := $(OUT)/obj
CXX      := g++
CXXFLAGS := -O3 -DNDEBUG -std=c++14 -w -I$(REF)/includes -I$(MPIROOT)/include
LIBDIR   := $(abspath $(OUT))/mpilib
LDFLAGS  := -L$(LIBDIR) -Wl,-rpath,$(LIBDIR) -lmpi

SRCS := src/BTEAngle/BTEAngle.cpp \
        src/BTEBand/BTEBand.cpp \
        src/BTEBoundaryCondition/BTEBoundaryCondition.cpp \
        src/BTEMesh/BTEMesh.cpp src/BTEMesh/COMSOL.cpp src/BTEMesh/oneD.cpp \
        src/BTEMesh/Distribute.cpp src/BTEMesh/ReadinMesh.cpp \
        src/StaticBTESynthetic/StaticBTESynthetic.cpp \
        src/StaticBTESynthetic/StaticBTESyntheticSolver.cpp \
        src/StaticFourier/StaticFourier.cpp \
        src/TransientBTE/transient.cpp src/TransientBTE/TransientSolver.cpp \
        src/Solution/Solution.cpp \
        src/utility/utility.cpp

OBJS := $(patsubst %.cpp,$(OBJ)/%.o,$(notdir $(SRCS)))

vpath %.cpp $(sort $(dir $(addprefix $(REF)/,$(SRCS))))

all: $(OUT)/ref_dump $(OUT)/BTE_CPU

$(LIBDIR)/.stamp:
	mkdir -p $(LIBDIR)
	for f in libmpi.so.12 libgfortran.so.4 libquadmath.so.0 libgomp.so.1 libgcc_s.so.1; do \
	  ln -sf $(MPIROOT)/lib/$$f $(LIBDIR)/$$f; done
	ln -sf $(MPIROOT)/lib/libmpi.so.12 $(LIBDIR)/libmpi.so
	touch $@

$(OBJ)/%.o: %.cpp
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ)/BTEMain.o: $(REF)/src/BTEMain.cpp
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJ)/ref_dump.o: oracle/ref_dump.cpp
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/ref_dump: $(OBJS) $(OBJ)/ref_dump.o $(LIBDIR)/.stamp
	$(CXX) -o $@ $(OBJ)/ref_dump.o $(OBJS) $(LDFLAGS)

$(OUT)/BTE_CPU: $(OBJS) $(OBJ)/BTEMain.o $(LIBDIR)/.stamp
	$(CXX) -o $@ $(OBJ)/BTEMain.o $(OBJS) $(LDFLAGS)

clean:
	rm -rf $(OUT)

.PHONY: all clean
