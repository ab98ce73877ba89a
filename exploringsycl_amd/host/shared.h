/* Constants and helpers shared by the host driver and the kernel adapter.
 * The numeric values of the face, field and coefficient identifiers are the
 * reference's (TeaLeaf/shared.h:28-51) because they cross the
 * kernel_interface boundary. */
#pragma once
#ifndef TEA_SHARED_H
#define TEA_SHARED_H

#include <stdarg.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
#include <cmath>
#endif

struct Settings;
/* Settings carries three profile pointers for layout only; this host times
 * with its own clock and leaves them null. */
struct Profile;

void initialise_log(struct Settings* settings);
void print_to_log(struct Settings* settings, const char* format, ...);
void print_and_log(struct Settings* settings, const char* format, ...);
void die(int lineNum, const char* file, const char* format, ...);

#define MASTER 0

/* Faces of a chunk; a neighbour of EXTERNAL_FACE is the domain boundary. */
#define NUM_FACES 4
#define CHUNK_LEFT 0
#define CHUNK_RIGHT 1
#define CHUNK_BOTTOM 2
#define CHUNK_TOP 3
#define EXTERNAL_FACE -1

/* Fields that halo updates can be requested for. */
#define FIELD_DENSITY 0
#define FIELD_ENERGY0 1
#define FIELD_ENERGY1 2
#define FIELD_U 3
#define FIELD_P 4
#define FIELD_SD 5

/* Conductivity models: K = density, or K = 1/density. */
#define CONDUCTIVITY 1
#define RECIP_CONDUCTIVITY 2

#define CG_ITERS_FOR_EIGENVALUES 20
#define ERROR_SWITCH_MAX 1.0

#define MIN(a, b) ((a < b) ? a : b)
#define MAX(a, b) ((a > b) ? a : b)
#define strmatch(a, b) (strcmp(a, b) == 0)

#endif
