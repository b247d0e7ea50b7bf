// Reference include path of the MPI communicator.  This build drives one
// GPU per process and has no MPI layer: a generator includes this header
// but only uses comm::mpi_t under __MPI__, which is never defined here.
#pragma once
#ifdef __MPI__
#error "this build has no MPI communicator"
#endif
