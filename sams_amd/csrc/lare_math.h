// Field names of the one-kernel-per-reference-kernel set (fields as
// lare::Arr::f[]) plus the shared numeric helpers.
#pragma once

#include "lare_dev.h"
#include "lare_num.h"

#define RHO A.f[0]
#define EE A.f[1]
#define EI A.f[2]
#define VX A.f[3]
#define VY A.f[4]
#define VZ A.f[5]
#define BX A.f[6]
#define BY A.f[7]
#define BZ A.f[8]
#define VX1 A.f[9]
#define VY1 A.f[10]
#define VZ1 A.f[11]
#define DM A.f[12]

// Field pointer by a launch-uniform index.  A switch over constant indices
// keeps the selection in scalar registers; indexing the kernel-argument
// array with a variable would copy the whole argument block to scratch.
__device__ __forceinline__ double *fsel(const lare::Arr &A, int i)
{
    switch (i) {
    case 0: return A.f[0];
    case 1: return A.f[1];
    case 2: return A.f[2];
    case 3: return A.f[3];
    case 4: return A.f[4];
    case 5: return A.f[5];
    case 6: return A.f[6];
    case 7: return A.f[7];
    case 8: return A.f[8];
    case 9: return A.f[9];
    case 10: return A.f[10];
    case 11: return A.f[11];
    default: return A.f[12];
    }
}

