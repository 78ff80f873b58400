/*
 * runner.h -- experiment harness (drop-in for the reference's
 * include/runner.h:6-17). Output lines keep the reference's formats
 * (runner.c:15-38, 52-67, 156-169). Thread counts are accepted for
 * compatibility; every column times the same GPU path.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces runner.c:8-43: prints "name: C0 C1 .. C(k-1)". */
void runner_test_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, const char *name, int k,
                          void (*f)(uint32_t, const uint32_t *, const uint32_t *, double *));

/* Replaces runner.c:45-72: prints "name: C seconds". */
void runner_test_ordering_time(uint32_t N, const uint32_t *V, const uint32_t *E, const char *name,
                               void (*f)(uint32_t, const uint32_t *, const uint32_t *, double *));

/* Replaces runner.c:85-176: best-of-it timing of setup'd f_seq / f_par with
 * Culberson chains of length k; prints "name_i colours t_seq t_tc1 .. t_tcn". */
void runner_test_full(uint32_t N, const uint32_t *V, const uint32_t *E,
                      void *(*setup)(uint32_t, uint32_t),
                      void (*cleanup)(void *),
                      void (*f_seq)(uint32_t, const uint32_t *, const uint32_t *, void *, void *, double *, int *),
                      void (*f_par)(uint32_t, const uint32_t *, const uint32_t *, void *, void *, double *, int *),
                      const char *s, int it, int k, int t_tot, int ntc, const int *nt);

#ifdef __cplusplus
}
#endif
