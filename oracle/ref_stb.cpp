// ref_stb.cpp: the vendored stb implementations the reference's backend sources call, instantiated for
// libnntc_ref.so (the reference's own main.cpp instantiates them for its binary: src/main.cpp:15-17).
// TEST INFRASTRUCTURE ONLY, compiled by oracle/Makefile against the headers where they lie.
#define STB_IMAGE_IMPLEMENTATION
#define STB_IMAGE_WRITE_IMPLEMENTATION
#define STB_IMAGE_RESIZE_IMPLEMENTATION
#include "stb_image.h"
#include "stb_image_write.h"
#include "stb_image_resize2.h"
