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

# --- the SDK with its GPU support on, and the HIP GPU plug-in ---------------
# The same sources compiled without DISABLE_CUDA_PHYSX (PX_SUPPORT_GPU_PHYSX)
# and with PX_PHYSX_GPU_STATIC, the static plug-in mode in which the SDK calls
# PxCreatePhysXGpu directly (physx/source/compiler/cmake/linux/PhysX.cmake:86-88,
# LowLevel.cmake:41-45). The plug-in (physx_amd/plugin/pxh_physxgpu.cpp, over
# libpxhip) becomes its own archive, and the test harness
# oracle/ref_gpu_shim.cpp links both into oracle/_ref/gpu/libpxrefgpu.so.
#
#   make -f oracle/ref.mk gpu -j8      (after make -C physx_amd/csrc)
GOUT     := $(OUT)/gpu
GOBJ     := $(GOUT)/obj
GCXXFLAGS := $(filter-out -DDISABLE_CUDA_PHYSX,$(CXXFLAGS)) -DPX_PHYSX_GPU_STATIC
GOBJS    := $(addprefix $(GOBJ)/,$(ALLSRC:.cpp=.o))
HIPINC   := -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -Iinclude

$(GOBJ)/%.o: $(SRC)/%.cpp
	@mkdir -p $(dir $@)
	@$(CXX) $(GCXXFLAGS) $(INC) -c $< -o $@

$(GOUT)/libphysx_refgpu.a: $(GOBJS)
	@rm -f $@
	@ar rcs $@ $(GOBJS)

$(GOUT)/pxh_physxgpu.o: physx_amd/plugin/pxh_physxgpu.cpp physx_amd/plugin/pxh_hip_rt.h include/pxhip.h
	@mkdir -p $(GOUT)
	$(CXX) $(GCXXFLAGS) $(INC) -Iinclude -c $< -o $@

# the plug-in's HIP runtime calls, compiled without any PhysX header
$(GOUT)/pxh_hip_rt.o: physx_amd/plugin/pxh_hip_rt.cpp physx_amd/plugin/pxh_hip_rt.h
	@mkdir -p $(GOUT)
	$(CXX) -std=c++14 -O2 -fPIC -fno-rtti -Wall $(HIPINC) -c $< -o $@

$(GOUT)/libPhysXGpuHip_static.a: $(GOUT)/pxh_physxgpu.o $(GOUT)/pxh_hip_rt.o
	@rm -f $@
	ar rcs $@ $^

$(GOUT)/libpxrefgpu.so: oracle/ref_gpu_shim.cpp $(GOUT)/libphysx_refgpu.a $(GOUT)/libPhysXGpuHip_static.a \
                        physx_amd/libpxhip.so
	$(CXX) -std=c++14 -O2 -fPIC -shared -fno-rtti -DNDEBUG -DPX_PHYSX_STATIC_LIB -DPX_PHYSX_GPU_STATIC -w $(INC) \
	  oracle/ref_gpu_shim.cpp -Wl,--whole-archive $(GOUT)/libphysx_refgpu.a $(GOUT)/libPhysXGpuHip_static.a \
	  -Wl,--no-whole-archive -Lphysx_amd -lpxhip -Wl,-rpath,'$$ORIGIN/../../../physx_amd' \
	  -L/opt/rocm/lib -lamdhip64 -lpthread -ldl -Wl,-Bsymbolic -Wl,--no-undefined -o $@

gpu: $(GOUT)/libpxrefgpu.so

clean:
	rm -rf $(OUT)

.PHONY: all gpu clean
