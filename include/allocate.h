/* Aligned host allocation; exits on failure (reference src/allocate.h, src/allocate.c:12-36). */
#ifndef NUSIF_ALLOCATE_H
#define NUSIF_ALLOCATE_H
#include <stddef.h>
#ifndef ARRAY_ALIGNMENT
#define ARRAY_ALIGNMENT 64
#endif
#ifdef __cplusplus
extern "C" {
#endif
extern void *allocate(size_t alignment, size_t bytesize);
#ifdef __cplusplus
}
#endif
#endif
