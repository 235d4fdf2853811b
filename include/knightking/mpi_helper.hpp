// mpi_helper.hpp -- the rank helpers KnightKing apps call
// (include/mpi_helper.hpp:34-46, and MPI_Instance from gemini), for one
// process driving one GPU: rank 0 of 1.  Lets app mains that construct an
// MPI_Instance and query ranks compile unchanged against this engine.
#pragma once

inline int get_mpi_rank() { return 0; }
inline int get_mpi_size() { return 1; }

class MPI_Instance {
public:
    MPI_Instance(int *, char ***) {}
};
