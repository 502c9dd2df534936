// sparc_host.h -- host-side helpers shared by the 2D and 3D C ABIs.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "sparc_device.h"

// Device buffers owned by a handle, freed together.
struct DevArena {
    std::vector<void *> ptrs;
    hipError_t alloc(double **p, size_t n)
    {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(double));
        if (e != hipSuccess) return e;
        ptrs.push_back(q);
        *p = (double *)q;
        return hipSuccess;
    }
    void release()
    {
        for (void *q : ptrs) (void)hipFree(q);
        ptrs.clear();
    }
};

// dbyd2 coefficients (dbyd2.f:85-162), float32 scalars widened.
sparc::DbyCoef make_coef(int N2);
// Thomas factors of the dbyd2 matrix of one ibc (dbyd2.f:151-195,
// mtridag.f:96-100): lowr, gam = uppr(j-1)/betad(j-1), 1/betad.
void make_tri(int N2, int ibc, std::vector<double> &lowr, std::vector<double> &gam,
              std::vector<double> &rbet);
// Upload a host vector into a new device buffer of the arena.
hipError_t upload(DevArena &ar, double **dst, const std::vector<double> &v, hipStream_t st);
