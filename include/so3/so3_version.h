/* so3_version.h -- version API (reference include/so3/so3_version.in.h:3-22). */
#ifndef SO3_VERSION_H
#define SO3_VERSION_H
#define SO3_VERSION_MAJOR 1
#define SO3_VERSION_MINOR 3
#define SO3_VERSION_PATCH 7
static inline const char *so3_version_string(void) { return "1.3.7"; }
static inline const char *so3_info(void) {
  return "package:\n"
         "  name: SO3\n"
         "  description: Fast and accurate Wigner transforms (MI355X FP64 build)\n"
         "  license: GPL-3\n"
         "  version: 1.3.7\n";
}
static inline int so3_version_major(void) { return SO3_VERSION_MAJOR; }
static inline int so3_version_minor(void) { return SO3_VERSION_MINOR; }
static inline int so3_version_patch(void) { return SO3_VERSION_PATCH; }
#endif
