// GL names used by the raycaster processors (values from the GL headers).
#pragma once
#include <cstdint>
typedef unsigned int GLenum;
typedef int GLint;
typedef unsigned char GLubyte;
typedef float GLfloat;
#define GL_NEAREST 0x2600
#define GL_LINEAR 0x2601
#define GL_RGBA 0x1908
#define GL_FLOAT 0x1406
#define GL_UNSIGNED_BYTE 0x1401
#define GL_COLOR_BUFFER_BIT 0x00004000
#define GL_DEPTH_BUFFER_BIT 0x00000100
#define GL_REPEAT 0x2901
#define GL_CLAMP_TO_EDGE 0x812F
inline void glClear(unsigned int) {}
