/*
 * Levelled logger (interface of include/log.h; behaviour of lib/log.c of the
 * reference: "YYYY-MM-DD HH:MM:SS.mmm LEVEL [subsys] message", INFO and
 * below to stdout, WARN and above to stderr, flushed per line).
 */
#define _POSIX_C_SOURCE 200809L
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <sys/time.h>
#include <time.h>

#include "log.h"

int log_threshold = LOG_INFO;

static const char *const NAMES[] = { "TRACE", "DEBUG", "INFO", "WARN",
	"ERROR", "FATAL" };

int log_init(void)
{
	const char *env = getenv("PHASE2_LOG");
	if (!env || !*env)
		return 0;
	for (int l = LOG_TRACE; l <= LOG_FATAL; l++)
		if (strcasecmp(env, NAMES[l]) == 0) {
			log_threshold = l;
			return 0;
		}
	return 0;
}

void log_emit(int level, const char *subsys, const char *file, int line,
	const char *fmt, ...)
{
	(void)file;
	(void)line;
	if (level < LOG_TRACE || level > LOG_FATAL)
		level = LOG_ERROR;
	FILE *out = level >= LOG_WARN ? stderr : stdout;
	struct timeval tv;
	gettimeofday(&tv, NULL);
	struct tm tm;
	localtime_r(&tv.tv_sec, &tm);
	char ts[32];
	strftime(ts, sizeof ts, "%Y-%m-%d %H:%M:%S", &tm);
	fprintf(out, "%s.%03d %-5s [%s] ", ts, (int)(tv.tv_usec / 1000),
		NAMES[level], subsys);
	va_list ap;
	va_start(ap, fmt);
	vfprintf(out, fmt, ap);
	va_end(ap);
	fputc('\n', out);
	fflush(out);
}
