/*
 * faf_serve.c -- request-serving loop (continuous batching); see below.
 */
#include "faf_internal.h"
