/* oracle_render.c -- CPU restatement of the per-pixel render path (TEST INFRASTRUCTURE). */
#include "oracle_render.h"
