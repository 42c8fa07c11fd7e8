# Builds the reference PhysX 5.6.1 CPU SDK (the CPU rigid-body pipeline that the
# GPU step must agree with) straight from the sources under /root/reference,
# without running the reference's own CMake/packman build system, and links it
# with the C-ABI harness oracle/ref_shim.cpp into oracle/_ref/libpxref.so.
#
# Nothing is copied out of /root/reference: object files, the archive and the
# shared library are written only to oracle/_ref/ (git-ignored; it travels to
# the GPU host so the CPU baseline can be timed there without the sources).
#
#   make -f oracle/ref.mk -j8          (from the repository root)
#
# Compile flags follow physx/source/compiler/cmake/linux/CMakeLists.txt:120-175
# (release configuration: NDEBUG, PX_SUPPORT_PVD=0, PX_SUPPORT_OMNI_PVD=0,
# DISABLE_CUDA_PHYSX, static libraries).

REF      ?= /root/reference/physx
SRC      := $(REF)/source
OUT      := oracle/_ref
OBJ      := $(OUT)/obj

CXX      ?= g++
CXXFLAGS := -std=c++14 -O3 -fPIC -D_GLIBCXX_USE_CXX11_ABI=1 -fno-rtti -fno-exceptions \
            -ffunction-sections -fdata-sections -fno-strict-aliasing -w \
            -DNDEBUG -DPX_SUPPORT_PVD=0 -DPX_SUPPORT_OMNI_PVD=0 -DDISABLE_CUDA_PHYSX \
            -DPX_PHYSX_STATIC_LIB -DPX_PUBLIC_RELEASE=1 -DPX_NVTX=0

INCDIRS  := include pvdruntime/include \
  source/common/include source/common/src source/filebuf/include \
  source/foundation/include \
  source/geomutils/include source/geomutils/src source/geomutils/src/ccd \
  source/geomutils/src/common source/geomutils/src/contact source/geomutils/src/convex \
  source/geomutils/src/distance source/geomutils/src/gjk source/geomutils/src/hf \
  source/geomutils/src/intersection source/geomutils/src/mesh source/geomutils/src/pcm \
  source/geomutils/src/sweep \
  source/lowlevel/api/include source/lowlevel/common/include \
  source/lowlevel/common/include/collision source/lowlevel/common/include/pipeline \
  source/lowlevel/common/include/utils source/lowlevel/software/include \
  source/lowlevelaabb/include source/lowlevelaabb/src \
  source/lowleveldynamics/include source/lowleveldynamics/shared source/lowleveldynamics/src \
  source/physx/src source/physx/src/internal/device source/physx/src/omnipvd \
  source/physxextensions/src source/physxextensions/src/serialization/Binary \
  source/physxextensions/src/serialization/File source/physxextensions/src/serialization/Xml \
  source/physxextensions/src/tet source/physxextensions/src/omnipvd \
  source/physxgpu/include source/physxmetadata/core/include \
  source/physxmetadata/extensions/include source/pvd/include source/pvd/src \
  source/scenequery/include source/scenequery/src \
  source/simulationcontroller/include source/simulationcontroller/src \
  source/immediatemode/include source/physxcooking/src source/cudamanager/include
INC      := $(addprefix -I$(REF)/,$(INCDIRS))

SRCDIRS  := foundation common geomutils lowlevel lowlevelaabb lowleveldynamics \
            simulationcontroller physx immediatemode scenequery physxextensions \
            task pvd physxcooking physxmetadata
ALLSRC   := $(shell cd $(SRC) && find $(SRCDIRS) -name '*.cpp' -not -path '*/windows/*' | sort)
OBJS     := $(addprefix $(OBJ)/,$(ALLSRC:.cpp=.o))

all: $(OUT)/libpxref.so

$(OBJ)/%.o: $(SRC)/%.cpp
	@mkdir -p $(dir $@)
	@$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OUT)/libphysx_ref.a: $(OBJS)
	@rm -f $@
	@ar rcs $@ $(OBJS)

$(OUT)/libpxref.so: oracle/ref_shim.cpp $(OUT)/libphysx_ref.a
	$(CXX) -std=c++14 -O2 -fPIC -shared -DNDEBUG -DPX_PHYSX_STATIC_LIB -w $(INC) \
	  oracle/ref_shim.cpp -Wl,--whole-archive $(OUT)/libphysx_ref.a -Wl,--no-whole-archive \
	  -lpthread -ldl -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
