/* Symbol visibility for libgrrt.so (stands where the reference's
 * CMake-generated export header stood; include/grrt/api.h includes it). */
#ifndef GRRT_EXPORT_H
#define GRRT_EXPORT_H
#if defined(GRRT_BUILDING_DLL)
#define GRRT_EXPORT __attribute__((visibility("default")))
#else
#define GRRT_EXPORT
#endif
#endif
