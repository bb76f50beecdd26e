#ifndef CONTAINER_OF_H
#define CONTAINER_OF_H

#include <stddef.h>

/* Pointer to the struct of type `type` whose member `member` is at ptr. */
#define container_of(ptr, type, member) \
	((type *)((char *)(ptr) - offsetof(type, member)))

#endif /* CONTAINER_OF_H */
