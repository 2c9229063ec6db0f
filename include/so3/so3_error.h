/*
 * so3_error.h -- fatal-error macros of the SO(3) C API.
 *
 * Same behaviour as the reference include/so3/so3_error.h:19-33: print
 * "ERROR: <text>." plus the location to stdout and exit(1).
 */
#ifndef SO3_ERROR
#define SO3_ERROR

#include <stdio.h>
#include <stdlib.h>

#define SO3_ERROR_GENERIC(comment)                                             \
  {                                                                            \
    printf("ERROR: %s.\n", comment);                                           \
    printf("ERROR: %s <%s> %s %s %s %d.\n", "Occurred in function",            \
           __func__, "of file", __FILE__, "on line", __LINE__);                \
    fflush(stdout);                                                            \
    exit(1);                                                                   \
  }

#define SO3_ERROR_MEM_ALLOC_CHECK(pointer)                                     \
  if ((pointer) == NULL) {                                                     \
    SO3_ERROR_GENERIC("Memory allocation failed")                              \
  }

#endif
