# HIPRaycaster module: a Voreen processor that runs the DVR ray march on an
# MI355X through libhipdvr.so.  Drop this directory into custommodules/ and
# enable VRN_MODULE_HIPRAYCASTER (cmake/commonconf.cmake:460-549 collects
# the module); template: modules/sample/sample.cmake.

# module class must reside in hipraycaster/hipraycastermodule.h + .cpp
SET(MOD_CORE_MODULECLASS HIPRaycasterModule)

SET(MOD_CORE_SOURCES
    ${MOD_DIR}/processors/hipraycaster.cpp
)

SET(MOD_CORE_HEADERS
    ${MOD_DIR}/processors/hipraycaster.h
)

# libhipdvr.so and its C-ABI header (built by voreen_amd/build.py)
SET(HIPDVR_ROOT "${MOD_DIR}/../.." CACHE PATH "root of the hipdvr checkout")
SET(MOD_INCLUDE_DIRECTORIES ${HIPDVR_ROOT}/include)
SET(MOD_LIBRARIES ${HIPDVR_ROOT}/voreen_amd/libhipdvr.so)
