/* num_defs.h -- numeric types of the rt engine plugin API
 * (same names and meaning as common-lib/num_defs.h of amirych/rt). */
#ifndef NUM_DEFS_H
#define NUM_DEFS_H

typedef double real_t;                  /* FP64 throughout */
#define RT_NUM_DOUBLE

#define RT_PI 3.14159265359             /* the reference's value of pi */
#define deg2rad RT_PI/180.0             /* unparenthesised, as callers rely on its expansion */

typedef char beam_flag_t;               /* 1: ray alive, 0: lost */

#endif /* NUM_DEFS_H */
