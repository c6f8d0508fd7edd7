// Process runtime of the Khuzdul application API on one MI355X.
//
// The reference initialises MPI here (include/distributed_sys.h:24-70,
// src/distributed_sys.cc:12-41) and its apps call MPI_Barrier directly
// (e.g. kautomine/triangle.cc:203).  This build runs as exactly one process
// on one GPU, so the few MPI names the apps use are provided as single-rank
// definitions (a barrier is a no-op, a reduction is a copy) by
// libkhuzdul_amd.so; no MPI library is needed.
#ifndef DISTRIBUTED_SYS_H
#define DISTRIBUTED_SYS_H

#include <stdint.h>
#include <string.h>

typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Op;
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#define MPI_CHAR ((MPI_Datatype)1)
#define MPI_INT ((MPI_Datatype)4)
#define MPI_UNSIGNED ((MPI_Datatype)5)
#define MPI_LONG_LONG ((MPI_Datatype)8)
#define MPI_UNSIGNED_LONG ((MPI_Datatype)9)
#define MPI_UNSIGNED_LONG_LONG ((MPI_Datatype)10)
#define MPI_DOUBLE ((MPI_Datatype)11)
#define MPI_FLOAT ((MPI_Datatype)12)
#define MPI_SUM ((MPI_Op)1)
#define MPI_MAX ((MPI_Op)2)
#define MPI_MIN ((MPI_Op)3)
#define MPI_SUCCESS 0

extern "C" {
int MPI_Barrier(MPI_Comm comm);
int MPI_Allreduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype type, MPI_Op op, MPI_Comm comm);
int MPI_Comm_rank(MPI_Comm comm, int* rank);
int MPI_Comm_size(MPI_Comm comm, int* size);
}

class DistributedSys {
   private:
    static DistributedSys* instance_;
    int node_id_ = 0;
    int num_nodes_ = 1;
    DistributedSys() {}

   public:
    static void init_distributed_sys();
    static void finalize_distributed_sys();
    static DistributedSys* get_instance();
    inline int get_node_id() { return node_id_; }
    inline int get_num_nodes() { return num_nodes_; }
    inline bool is_master_node() { return node_id_ == 0; }
    template <typename T>
    static MPI_Datatype get_mpi_data_type() {
        return sizeof(T) == 8 ? MPI_UNSIGNED_LONG_LONG : MPI_UNSIGNED;
    }
};

#endif
