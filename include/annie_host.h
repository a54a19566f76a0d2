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

#ifdef __cplusplus
}
#endif

#endif /* ANNIE_HOST_H */
