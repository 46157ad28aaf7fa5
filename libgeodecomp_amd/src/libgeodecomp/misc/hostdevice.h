// __host__ / __device__ for code shared by host compilers and hipcc.
//
// User cells that run on the GPU through the generic kernel family mark
// update() (and what it calls) __host__ __device__, as CUDA cells of the
// reference do; the same cell is compiled by g++ for the CPU oracle and by
// plain C++ host code, where the two attributes expand to nothing (the
// reference's src/libgeodecomp/misc/cudautil.h:26-35 does the same for
// non-CUDA builds).  Under hipcc (clang in HIP mode, __HIP__ defined) the
// compiler's own definitions are used.
#ifndef LIBGEODECOMP_MISC_HOSTDEVICE_H
#define LIBGEODECOMP_MISC_HOSTDEVICE_H

#ifndef __HIP__
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#endif

#endif
