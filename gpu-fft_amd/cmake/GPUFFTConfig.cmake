# CMake package of the MI355X build: `find_package(GPUFFT CONFIG REQUIRED)`
# then `target_link_libraries(app PRIVATE GPUFFT::fft)`, the target name of
# the reference's package (cmake/Config.cmake.in, src/CMakeLists.txt:54-85).
# Works from the source tree (GPUFFT_DIR=<repo>/gpu-fft_amd/cmake) and from
# an installed prefix (`make -C gpu-fft_amd install PREFIX=...` puts this file
# in <prefix>/lib/cmake/GPUFFT-1.0).  GPUFFT::fft carries the include
# directory of gpufft_amd.h and gpufft/*.cuh and links hip::host.
include(CMakeFindDependencyMacro)
find_dependency(hip CONFIG HINTS "$ENV{ROCM_PATH}" /opt/rocm)

get_filename_component(_gpufft_here "${CMAKE_CURRENT_LIST_DIR}" ABSOLUTE)
get_filename_component(_gpufft_prefix "${_gpufft_here}/../../.." ABSOLUTE)
if(EXISTS "${_gpufft_prefix}/lib/libgpufft_amd.so"
   AND EXISTS "${_gpufft_prefix}/include/GPUFFT-1.0/gpufft_amd.h")
  set(_gpufft_lib "${_gpufft_prefix}/lib/libgpufft_amd.so")
  set(_gpufft_inc "${_gpufft_prefix}/include/GPUFFT-1.0")
else()
  get_filename_component(_gpufft_lib "${_gpufft_here}/../lib/libgpufft_amd.so" ABSOLUTE)
  get_filename_component(_gpufft_inc "${_gpufft_here}/../../include" ABSOLUTE)
endif()
if(NOT EXISTS "${_gpufft_lib}")
  set(GPUFFT_FOUND FALSE)
  set(GPUFFT_NOT_FOUND_MESSAGE "libgpufft_amd.so not found at ${_gpufft_lib} (build it with make -C gpu-fft_amd)")
  return()
endif()

if(NOT TARGET GPUFFT::fft)
  add_library(GPUFFT::fft SHARED IMPORTED)
  set_target_properties(GPUFFT::fft PROPERTIES
    IMPORTED_LOCATION "${_gpufft_lib}"
    INTERFACE_INCLUDE_DIRECTORIES "${_gpufft_inc}"
    INTERFACE_LINK_LIBRARIES hip::host)
endif()
set(GPUFFT_INCLUDE_DIRS "${_gpufft_inc}")
set(GPUFFT_LIBRARIES GPUFFT::fft)
