// clg_packets.h -- 16-byte packet views of the flat BLAS arrays: a fermion field (or an 18-real gauge
// field) is one contiguous array of complex numbers; a lane loads one packet, NC complex numbers,
// and the kernels compute on them as doubles.
#pragma once

#include "clg_internal.h"

namespace clgd {

// ---- 16-byte packets: NC complex numbers per lane and load, values handled as doubles
enum { PK_D = 0, PK_F4 = 1, PK_F2 = 2 };
template <int P> struct Pk;
template <> struct Pk<PK_D> {
    typedef double2 T;
    static constexpr int NC = 1;
    __device__ static void cvt(const T& a, double* v) { v[0] = a.x; v[1] = a.y; }
    __device__ static void ld(const T* p, long i, double* v) {
        const double2 a = p[i];
        v[0] = a.x; v[1] = a.y;
    }
    __device__ static void st(T* p, long i, double* v) { p[i] = make_double2(v[0], v[1]); }
};
template <> struct Pk<PK_F4> {
    typedef float4 T;
    static constexpr int NC = 2;
    __device__ static void cvt(const T& a, double* v) { v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; }
    __device__ static void ld(const T* p, long i, double* v) {
        const float4 a = p[i];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
    // stores the FP32-rounded value and leaves it in v (so norms see what was stored)
    __device__ static void st(T* p, long i, double* v) {
        const float4 a = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
        p[i] = a;
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
};
template <> struct Pk<PK_F2> {
    typedef float2 T;
    static constexpr int NC = 1;
    __device__ static void cvt(const T& a, double* v) { v[0] = a.x; v[1] = a.y; }
    __device__ static void ld(const T* p, long i, double* v) {
        const float2 a = p[i];
        v[0] = a.x; v[1] = a.y;
    }
    __device__ static void st(T* p, long i, double* v) {
        const float2 a = make_float2((float)v[0], (float)v[1]);
        p[i] = a;
        v[0] = a.x; v[1] = a.y;
    }
};

}  // namespace clgd
