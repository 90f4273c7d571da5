/*
 * ftcrypto.h -- C interface of the MI355X (gfx950) batched Ed25519 library.
 *
 * Declaration-compatible with the reference interface (ftcrypto.h:1-437 of
 * AUSTRAC/libftcrypto): the same 38 functions, types, sizes and error codes,
 * so a caller compiled against the reference header links against
 * libftcrypto.so unchanged.  Each declaration names the reference line it
 * replaces.  Deviations (all documented in the repository's design notes):
 * device errors return FT_ERR_INTERNAL_LIBRARY_ERROR / FT_ERR_OUT_OF_MEMORY
 * instead of exiting the process; out-of-range indices return
 * FT_ERR_OUT_OF_BOUNDS; ft_neg computes -P (the reference returns
 * -P + (0,-1)); ft_array_get_range returns ceil((stop-start)/step) elements
 * as the reference header documents.
 */
#ifndef FT_GE25519_ARRAY
#define FT_GE25519_ARRAY

#include <stdint.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FT_ED25519_SCALARBYTES 32                       /* ftcrypto.h:19 */
#define FT_FOLDED_POINT_BYTES 32                        /* ftcrypto.h:25 */
#define FT_EXTENDED_POINT_BYTES 160                     /* ftcrypto.h:31 (sizeof(ge25519_p3)) */
#define FT_AFFINE_POINT_BYTES (2 * FT_ED25519_SCALARBYTES) /* ftcrypto.h:38 */

struct ft_ge25519_array_s;                              /* ftcrypto.h:44-46 */
typedef struct ft_ge25519_array_s *ft_ge25519_array;
typedef struct ft_ge25519_array_s const *const_ft_ge25519_array;

typedef int ft_error;                                   /* ftcrypto.h:51-63 */
#define FT_ERR_NO_ERROR 0
#define FT_ERR_OUT_OF_MEMORY 1
#define FT_ERR_SIZE_MISMATCH 2
#define FT_ERR_BAD_SIZE 3
#define FT_ERR_OUT_OF_BOUNDS 4
#define FT_ERR_INVALID_POINT 5
#define FT_ERR_SODIUM_INIT 6
#define FT_ERR_INSUFFICIENT_SPACE 7
#define FT_ERR_BAD_CIPHERTEXT 8
#define FT_ERR_INTERNAL_LIBRARY_ERROR 9
#define FT_ERR_INSUFFICIENT_DEVICE_OR_DRIVER 10
#define FT_ERR_MAX_ERR 10

const char *ft_error_str(ft_error err);                 /* ftcrypto.h:65 */
ft_error ft_crypto_init(void);                          /* ftcrypto.h:71 */
ft_error ft_crypto_device_memory(size_t *free_mem, size_t *total_mem);   /* ftcrypto.h:76 */

ft_error ft_array_init_scalar(ft_ge25519_array *array, size_t n_elts,    /* ftcrypto.h:86-88 */
                              const uint8_t *scalar);
ft_error ft_array_init_point(ft_ge25519_array *array, size_t n_elts,     /* ftcrypto.h:95-97 */
                             const_ft_ge25519_array src, size_t idx);
ft_error ft_array_get_subset(ft_ge25519_array *array, const size_t *idxs, size_t n_idxs,  /* :105-108 */
                             const_ft_ge25519_array src);
ft_error ft_array_get_range(ft_ge25519_array *array, size_t start, size_t stop, size_t step, /* :116-119 */
                            const_ft_ge25519_array src);
ft_error ft_array_assign_to_subset(ft_ge25519_array array, const size_t *idxs, size_t n_idxs, /* :124-127 */
                                   const_ft_ge25519_array src);
ft_error ft_array_assign_zero_to_subset(ft_ge25519_array array, const size_t *idxs,          /* :132-134 */
                                        size_t n_idxs);
ft_error ft_array_assign_to_range(ft_ge25519_array array, size_t start, size_t stop,         /* :139-142 */
                                  size_t step, const_ft_ge25519_array src);
ft_error ft_array_delete_subset(ft_ge25519_array array, const size_t *idxs, size_t n_idxs); /* :147-149 */
ft_error ft_array_delete_range(ft_ge25519_array array, size_t start, size_t stop,           /* :154-156 */
                               size_t step);
ft_error ft_array_to_bytes_folded(uint8_t *mem, size_t n_bytes,          /* ftcrypto.h:167-169 */
                                  const_ft_ge25519_array array);
ft_error ft_array_to_bytes_affine(uint8_t *mem, size_t n_bytes,          /* ftcrypto.h:181-183 */
                                  const_ft_ge25519_array array);
ft_error ft_array_to_bytes(uint8_t *mem, size_t n_bytes,                 /* ftcrypto.h:190-192 */
                           const_ft_ge25519_array array);
ft_error ft_array_from_bytes_folded(int64_t *res, ft_ge25519_array *array,  /* ftcrypto.h:210-213 */
                                    const uint8_t *mem, size_t n_bytes);
ft_error ft_array_from_bytes_affine(ft_ge25519_array *array,             /* ftcrypto.h:229-231 */
                                    const uint8_t *mem, size_t n_bytes);
ft_error ft_array_from_bytes(ft_ge25519_array *array,                    /* ftcrypto.h:247-249 */
                             const uint8_t *mem, size_t n_bytes);
ft_error ft_array_validate(int64_t *res, const_ft_ge25519_array array);  /* ftcrypto.h:261-263 */
ft_error ft_array_from_scalars(ft_ge25519_array *array,                  /* ftcrypto.h:275-277 */
                               const uint8_t *scalars, size_t n_scalars);
ft_error ft_array_from_small_scalars(ft_ge25519_array *array,            /* ftcrypto.h:282-284 */
                                     const uint64_t *small_scalars, size_t n_scalars);
void ft_array_free(ft_ge25519_array array);                              /* ftcrypto.h:289 */
size_t ft_array_get_length(const_ft_ge25519_array array);                /* ftcrypto.h:294 */
ft_error ft_array_set_length(ft_ge25519_array array, size_t newlen);     /* ftcrypto.h:308 */
ft_error ft_array_mux(ft_ge25519_array iftrue, const int64_t *cond,      /* ftcrypto.h:315-318 */
                      const_ft_ge25519_array iffalse);
ft_error ft_array_contains(int64_t *res, const_ft_ge25519_array array1,  /* ftcrypto.h:326-329 */
                           const_ft_ge25519_array array2);
ft_error ft_add(ft_ge25519_array res, const_ft_ge25519_array summand);   /* ftcrypto.h:335-337 */
ft_error ft_reduce_isum(ft_ge25519_array res, const size_t *idxs,        /* ftcrypto.h:345-348 */
                        size_t n_idxs, const_ft_ge25519_array summand);
ft_error ft_sub(ft_ge25519_array res, const_ft_ge25519_array summand);   /* ftcrypto.h:354-356 */
void ft_scale(ft_ge25519_array array, const uint8_t *scalar);            /* ftcrypto.h:365-367 */
void ft_neg(ft_ge25519_array array);                                     /* ftcrypto.h:372 */
ft_error ft_mul(ft_ge25519_array array, const uint8_t *scalars,          /* ftcrypto.h:383-385 */
                size_t n_scalars);
void ft_index(int64_t *res, const_ft_ge25519_array array);               /* ftcrypto.h:391-393 */
ft_error ft_equal(int64_t *res, const_ft_ge25519_array array1,           /* ftcrypto.h:402-405 */
                  const_ft_ge25519_array array2);
ft_error ft_not_equal(int64_t *res, const_ft_ge25519_array array1,       /* ftcrypto.h:414-417 */
                      const_ft_ge25519_array array2);
ft_error ft_array_prescan(ft_ge25519_array array);                       /* ftcrypto.h:425 */
ft_error ft_array_scan(ft_ge25519_array array);                          /* ftcrypto.h:432 */
ft_error ft_array_reduce(ft_ge25519_array array);                        /* ftcrypto.h:435 */

#ifdef __cplusplus
}
#endif

#endif
