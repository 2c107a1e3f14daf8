/* oracle_render.h -- CPU restatement of the per-pixel render path (TEST INFRASTRUCTURE). */
#ifndef ORF_ORACLE_RENDER_H
#define ORF_ORACLE_RENDER_H
#include "orf_oracle.h"
#endif
