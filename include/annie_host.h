/*
 * annie_host.h -- C ABI of the test-protocol host that drives the gfx950
 * engine (the reference's caller side, src/annie.h + src/file.h), so that
 * bindings in other languages can run a whole test with one call.
 */
#ifndef ANNIE_HOST_H
#define ANNIE_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int is_euler;               /* Euler test (Proth's theorem for k*2^n+1) instead of Fermat */
  int32_t a_ref;              /* Fermat base, 0 = default 2 (src/annie.h:249-254) */
  int device;                 /* GPU ordinal */
  int gmp_threshold_digits;   /* below this many digits use the GMP transform (reference: 10000) */
  int min_ln;                 /* smallest accepted log2 transform length (reference: 10) */
  const char* workdir;        /* directory of checkpoint / log / result files, NULL = none written */
  double checkpoint_seconds;  /* checkpoint period (reference: 600 s) */
  long quit_after;            /* >= 0: stop (checkpoint + Aborted) after that many loop iterations */
  long corrupt_at;            /* >= 0: corrupt Reg0 at that loop iteration (GL check must fail) */
  int verbose;                /* print progress / result lines */
  const char* expression;     /* N as typed, named in messages when at most 40 characters
                                 (src/annie.h:219); NULL = the decimal value */
} annie_check_opts;

typedef struct {
  int status;                 /* 0 Success, 1 Failed (GL mismatch), 2 Aborted, 3 factor found,
                                 4 decided without the loop (<= 15 digits or even), -1 error */
  int is_prp;
  uint64_t res64;
  int gl_ok;
  int32_t a;                  /* base actually used */
  int resumed;                /* 1 if the run resumed from a checkpoint */
  long iterations;            /* loop iterations executed by this call */
  double seconds;
  char message[1024];         /* result line (src/annie.h:412-426) or error text */
} annie_check_result;

/* Replaces annie::check (src/annie.h:217-430). */
int annie_check(const uint64_t* N, size_t nlimbs, const annie_check_opts* opts, annie_check_result* res);

/* Sets the quit flag the hot loop polls (src/annie.h:57-69; SIGINT/SIGTERM). */
void annie_request_quit(void);
void annie_clear_quit(void);

/* Checkpoint file name for N (src/annie.h:288-291). */
int annie_checkpoint_name(const uint64_t* N, size_t nlimbs, char* buf, size_t cap);

/* CRC-32 of the checkpoint files (src/file.h:56-90): running value after buf. */
uint32_t annie_crc32(uint32_t crc, const void* buf, size_t len);

/* ---- Batched and served tests (no counterpart in the reference, which runs
 * one test per process; these drive many tests of a length class through one
 * engine batch with continuous admission). */

typedef struct {
  const uint64_t* N;          /* little-endian 64-bit limbs */
  size_t nlimbs;
  int is_euler;               /* Euler (Proth) test instead of Fermat */
  int32_t a_ref;              /* Fermat base, 0 = default 2 */
  double arrival;             /* seconds after the start of the run (serving) */
} annie_batch_test;

typedef struct {
  int engine;                 /* 0: gfx950 engine; 1: exact GMP transforms on the CPU (tests) */
  int device;                 /* GPU ordinal */
  size_t max_slots;           /* slots per length class, 0 = one per test of the class */
  double hbm_budget_mb;       /* device memory for all class batches, 0 = 64 GiB */
  int realtime;               /* 1: admit a test only once its arrival time has passed */
  double time_limit;          /* > 0: stop after that many seconds, not counting batch set-up
                                 (unfinished tests: status 2) */
  int verbose;                /* print each result line on stderr as it is decided */
  int graphs;                 /* 1: batched squarings replay captured HIP graphs */
  int flush_ticks;            /* deferred Gerbicz-Li products taken every that many ticks, 0 = 16 */
  int run_steps;              /* > 0: a class with at least run_min_active running tests takes up to
                                 that many squarings per test in one launch (squaring runs, lengths
                                 2^9 .. 2^13); 0: one squaring per tick */
  int run_min_active;         /* running tests a class needs for squaring runs, 0 = 1 */
  int host_threads;           /* host worker threads driving the classes (class i to worker i mod W),
                                 0 = min(classes, 16); one when graphs is set */
  int row_lines, col_cols;    /* NTT launch schedule of every class whose transform length it fits
                                 (annie_hip_set_schedule in annie_hip.h); 0 = default */
} annie_batch_opts;

typedef struct {
  int status;                 /* as annie_check_result.status */
  int is_prp;
  uint64_t res64;
  int gl_ok;
  int32_t a;
  long iterations;            /* loop iterations done */
  double arrival, admitted, finished;  /* seconds after the start of the run */
  char message[256];
} annie_batch_result;

typedef struct {
  double seconds;             /* wall time of the run */
  double squarings;           /* square_mul operations issued, summed over tests */
  int classes;                /* length classes used */
  size_t slots;               /* slots allocated over all classes */
  double setup_seconds;       /* part of `seconds` spent creating batches and loading moduli */
} annie_batch_stats;

/* Runs every test to completion (or to time_limit); results[i] belongs to tests[i]. */
int annie_batch_run(const annie_batch_test* tests, size_t count, const annie_batch_opts* opts,
                    annie_batch_result* results, annie_batch_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* ANNIE_HOST_H */
