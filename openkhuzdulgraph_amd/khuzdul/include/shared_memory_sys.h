// Host-thread limits kept for source compatibility (include/shared_memory_sys.h:6-7).
// The MI355X engine runs no host compute threads; counting happens on the GPU.
#ifndef SHARED_MEMORY_SYS_H
#define SHARED_MEMORY_SYS_H

#define MAX_NUM_THREADS 64
#define MAX_NUM_SOCKETS 8

class SharedMemorySys {
   public:
    static void init_shared_memory_sys() {}
    static int get_num_threads();
};

#endif
