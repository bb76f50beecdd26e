#ifndef LOG_H
#define LOG_H

/*
 * Levelled logging with the phase2 interface (replaces include/log.h of the
 * reference): log_info/warn/error/fatal always compile; log_trace and
 * log_debug only with -DDEBUG.  The threshold comes from PHASE2_LOG
 * (trace, debug, info, warn, error, fatal; default info).  Messages at
 * warn and above go to stderr, the rest to stdout.  Define LOG_SUBSYS
 * before including to tag messages.
 */

#ifndef LOG_SUBSYS
#define LOG_SUBSYS "?"
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum log_level {
	LOG_TRACE = 0,
	LOG_DEBUG = 1,
	LOG_INFO = 2,
	LOG_WARN = 3,
	LOG_ERROR = 4,
	LOG_FATAL = 5,
};

extern int log_threshold;

int log_init(void);
void log_emit(int level, const char *subsys, const char *file, int line,
	const char *fmt, ...) __attribute__((format(printf, 5, 6)));

#define log_at(lvl, ...)                                                    \
	do {                                                                \
		if ((lvl) >= log_threshold)                                 \
			log_emit((lvl), LOG_SUBSYS, __FILE__, __LINE__,     \
				__VA_ARGS__);                               \
	} while (0)

#ifdef DEBUG
#define log_trace(...) log_at(LOG_TRACE, __VA_ARGS__)
#define log_debug(...) log_at(LOG_DEBUG, __VA_ARGS__)
#else
#define log_trace(...) ((void)0)
#define log_debug(...) ((void)0)
#endif
#define log_info(...) log_at(LOG_INFO, __VA_ARGS__)
#define log_warn(...) log_at(LOG_WARN, __VA_ARGS__)
#define log_error(...) log_at(LOG_ERROR, __VA_ARGS__)
#define log_fatal(...) log_at(LOG_FATAL, __VA_ARGS__)

#ifdef __cplusplus
}
#endif

#endif /* LOG_H */
