#ifndef C23_COMPAT_H
#define C23_COMPAT_H

/* C23 spellings used by the algorithm sources, for a C11 compiler. */
#include <stdbool.h>
#ifndef __cplusplus
#define nullptr ((void *)0)
#endif
#define unreachable() (__builtin_unreachable())
#ifndef static_assert
#define static_assert _Static_assert
#endif

#endif /* C23_COMPAT_H */
