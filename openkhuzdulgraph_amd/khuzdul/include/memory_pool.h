// Present for source compatibility: the reference's memory pools
// (include/memory_pool.h) are used by no app and have no role here.
#ifndef MEMORY_POOL_H
#define MEMORY_POOL_H
#endif
