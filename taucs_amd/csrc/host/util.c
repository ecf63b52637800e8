/*
 * Host utilities behind the TAUCS ABI: logging, timers, allocation stubs with
 * failure injection.
 *
 *   taucs_logfile / taucs_printf          src/taucs_logging.c:32-79
 *   taucs_wtime / taucs_ctime             src/taucs_timer.c:250-297
 *   allocation stubs + induced failure    src/taucs_malloc.c:127,166-302
 *
 * The logging sink is process-global like the reference's, but guarded by a
 * mutex so the serving loop may log from several threads.  The wall clock is
 * CLOCK_MONOTONIC (ns resolution) instead of the reference's 1 ms ftime.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/resource.h>
#include <time.h>

#include "internal.h"

static pthread_mutex_t log_mutex = PTHREAD_MUTEX_INITIALIZER;
static FILE* log_sink = NULL;
static int   log_mode = 0; /* 0 none, 1 stdout, 2 stderr, 3 file */

void taucs_logfile(char* file_prefix)
{
  pthread_mutex_lock(&log_mutex);
  if (log_mode == 3 && log_sink) fclose(log_sink);
  log_sink = NULL;
  log_mode = 0;
  if (!file_prefix || !strcmp(file_prefix, "none")) {
    log_mode = 0;
  } else if (!strcmp(file_prefix, "stdout")) {
    log_mode = 1;
  } else if (!strcmp(file_prefix, "stderr")) {
    log_mode = 2;
  } else {
    log_sink = fopen(file_prefix, "w");
    log_mode = log_sink ? 3 : 0;
  }
  pthread_mutex_unlock(&log_mutex);
}

int taucs_printf(char* fmt, ...)
{
  va_list ap;
  FILE* f;
  pthread_mutex_lock(&log_mutex);
  f = log_mode == 1 ? stdout : log_mode == 2 ? stderr : log_mode == 3 ? log_sink : NULL;
  if (f) {
    va_start(ap, fmt);
    vfprintf(f, fmt, ap);
    va_end(ap);
    fflush(f);
  }
  pthread_mutex_unlock(&log_mutex);
  return 0;
}

double taucs_wtime(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double taucs_ctime(void)
{
  struct rusage r;
  getrusage(RUSAGE_SELF, &r);
  return (double)r.ru_utime.tv_sec + 1e-6 * (double)r.ru_utime.tv_usec
       + (double)r.ru_stime.tv_sec + 1e-6 * (double)r.ru_stime.tv_usec;
}

/* ---- allocation stubs ---------------------------------------------------
 * Every host allocation of the path goes through these, so a test can make
 * the i-th one fail (the model is progs/xtest_allocations.c:62-210). */

static int alloc_attempts = 0;
static int alloc_fail_at  = -1;
static pthread_mutex_t alloc_mutex = PTHREAD_MUTEX_INITIALIZER;

static int alloc_should_fail(void)
{
  int fail;
  pthread_mutex_lock(&alloc_mutex);
  fail = (alloc_fail_at >= 0 && alloc_attempts == alloc_fail_at);
  alloc_attempts++;
  pthread_mutex_unlock(&alloc_mutex);
  return fail;
}

void taucs_allocation_induce_failure(int i)
{
  pthread_mutex_lock(&alloc_mutex);
  alloc_fail_at  = i;
  alloc_attempts = 0;
  pthread_mutex_unlock(&alloc_mutex);
}

int taucs_allocation_attempts(void) { return alloc_attempts; }

void* taucs_malloc_stub(size_t size)
{
  if (alloc_should_fail()) return NULL;
  return malloc(size ? size : 1);
}

void* taucs_calloc_stub(size_t nmemb, size_t size)
{
  if (alloc_should_fail()) return NULL;
  return calloc(nmemb ? nmemb : 1, size ? size : 1);
}

void* taucs_realloc_stub(void* ptr, size_t size)
{
  if (alloc_should_fail()) return NULL;
  return realloc(ptr, size ? size : 1);
}

void taucs_free_stub(void* ptr) { free(ptr); }

/* Device allocations share the same counter so a failure can be injected
   into the i-th hipMalloc as well. */
int tamd_alloc_should_fail(void) { return alloc_should_fail(); }
