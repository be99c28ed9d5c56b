/*
 * Shared constants of the OpenDBA DTW/DBA hot path: value-type codes, DTW step
 * codes and error codes. Plain C, no dependencies.
 */
#ifndef ODBA_TYPES_H
#define ODBA_TYPES_H

/* Value types the reference instantiates (openDBA.cu:121-143). The CLI type
 * "short" runs as ODBA_F32 with int16 binary input expanded to float
 * (openDBA.cu:140-143, cpu_utils.hpp:81-85). */
#define ODBA_I32 0 /* int                 */
#define ODBA_U32 1 /* unsigned int        */
#define ODBA_U64 2 /* unsigned long long  */
#define ODBA_F32 3 /* float               */
#define ODBA_F64 4 /* double              */

/* DTW step codes, identical to dtw.hpp:10-16. */
#define ODBA_NIL 255
#define ODBA_DIAGONAL 1
#define ODBA_RIGHT 2
#define ODBA_UP 3
#define ODBA_OPEN_RIGHT 4
#define ODBA_NIL_OPEN_RIGHT 5

/* Algorithm modes of performDBA (dba.hpp:30-32). */
#define ODBA_CLUSTER_ONLY 1
#define ODBA_CONSENSUS_ONLY 2
#define ODBA_CLUSTER_AND_CONSENSUS 3

/* Swath width the reference uses: maxThreadsPerBlock for the medoid search
 * (cuda_utils.hpp:42-57) and a fixed 1024 in DBAUpdate (dba.hpp:453-457). */
#define ODBA_REF_W 1024

/* Path-matrix policy of a DBA round (dba.hpp:523-548): the reference picks
 * stripe mode when free device memory is below the 1-byte-per-cell move
 * matrix plus 5%. */
#define ODBA_PATH_AUTO 0   /* reference rule against a stated free-memory figure */
#define ODBA_PATH_FULL 1   /* always the full move matrix                        */
#define ODBA_PATH_STRIPE 2 /* always stripe mode (store swath edges, recompute)  */

/* All-vs-all schedules (no counterpart in the reference, which launches one
 * threadblock per pair and swath). */
#define ODBA_SCHED_WAVE 0  /* one wave per pair, persistent grid               */
#define ODBA_SCHED_BANDS 1 /* bands of column strips, a workgroup per band     */
#define ODBA_SCHED_SMALL 2 /* small pairs (float, lengths 64..1024): a workgroup
                              of band_waves waves per pair, band_c (1, 2 or 4)
                              columns per lane, strips in lock step, ring =
                              steps per sub-block (16 or 32; 0 means 16)      */

typedef struct odba_schedule {
  int kind;       /* ODBA_SCHED_WAVE, ODBA_SCHED_BANDS or ODBA_SCHED_SMALL             */
  int order;      /* wave: pair packing order, 0 = second index descending then first   */
                  /* index descending, 1 = largest cell count first                     */
  int band_c;     /* bands: columns per lane, 1, 4 or 16 (strips of 64 * band_c columns) */
  int band_waves; /* bands: strips per band = waves per workgroup, 1..16                 */
  int ring;       /* bands: rows of the in-workgroup edge ring (power of two >= 128);    */
                  /* small: steps per sub-block (16 or 32, 0 means 16)                  */
} odba_schedule;

/* Return codes. 0 is success; 18..46 are the reference's exit codes
 * (exit_codes.hpp:4-32); values from 100 up are this library's own. */
#define ODBA_OK 0
#define ODBA_E_NOT_ENOUGH_SEQUENCES 32   /* exit_codes.hpp:18 */
#define ODBA_E_MEDOID 36                 /* exit_codes.hpp:23 */
#define ODBA_E_HIP 100          /* a HIP runtime call failed                    */
#define ODBA_E_INVALID 101      /* bad argument (null pointer, size, dtype ...) */
#define ODBA_E_NOMEM 102        /* device or host allocation failed             */
#define ODBA_E_TRACEBACK 103    /* traceback did not end at (0,0); the reference
                                   executes a trap here (dba.hpp:405-419)      */
#define ODBA_E_DIV_ZERO 104     /* integer centroid position with zero count;
                                   the reference divides by zero (dba.hpp:821) */
#define ODBA_E_NO_DEVICE 105    /* no HIP device or the device index is invalid */
#define ODBA_E_UNSUPPORTED 106  /* a mode this build rejects explicitly         */
#define ODBA_E_BUSY 107         /* serving: job not finished, or server stopping */

/* Serving job states (odba_server_wait). */
#define ODBA_JOB_QUEUED 0
#define ODBA_JOB_RUNNING 1
#define ODBA_JOB_DONE 2
#define ODBA_JOB_FAILED 3

#endif
