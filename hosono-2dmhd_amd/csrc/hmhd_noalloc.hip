// hmhd_noalloc.hip -- hmhd_cfl_dt_noalloc: the CFL time step of a state the
// caller keeps in device memory, computed without allocating anything.
//
// hmhd_cfl_dt reduces per-block maxima of |v| + c_f held in a buffer the
// engine allocates.  This variant keeps them in `scratch`, a file-scope
// device array of a fixed size (kNoallocBlocks blocks), so a caller that
// cannot call hipMalloc -- code under HIP graph capture, a tight serving
// loop -- can still compute dt.  The kernels and the arithmetic are those of
// hmhd_cfl_dt (src/cfl.f90:14-52): an exact maximum over bit patterns, then
// dt = cfl * min(dx / xmax, dy / ymax), so both give the same bits.
//
// Every call on one GPU shares that GPU's scratch array: calls that overlap
// on one GPU (other streams, other host threads) race and must be ordered
// by the caller.  No part of scratch needs clearing: each launch writes the
// partials of exactly the blocks its reduction then reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/hmhd.h"
#include "hmhd_internal.hpp"
#include "hmhd_math.hpp"

namespace hmhd {

constexpr int kNoallocBlocks = 1024;  // two 8-byte maxima per block: 16 KiB
static __device__ __align__(256) unsigned long long scratch[2 * kNoallocBlocks];

struct NoallocCfl {
    const double* U;  // U(8, mx, my), column-major, device memory
    const double* V;  // V(10, mx, my) or null
    int nx, ny, ngc, mx, hydro;
    double gamma, small_p, small_rho;
};

// per-block maxima over the interior (grid-stride), into part[2 b], part[2 b + 1]
__global__ __launch_bounds__(256) void k_cfl_noalloc(const NoallocCfl a, unsigned long long* __restrict__ part) {
    unsigned long long bx = 0, by = 0;
    const long long ncell = (long long)a.nx * a.ny;
    for (long long o = blockIdx.x * 256ll + threadIdx.x; o < ncell; o += (long long)gridDim.x * 256) {
        const int kj = (int)(o / a.nx), ki = (int)(o - (long long)kj * a.nx);
        const size_t c = (size_t)(kj + a.ngc) * a.mx + (ki + a.ngc);
        double d, vx, vy, p, b1, b2, b3;
        if (a.V) {
            const double* v = a.V + c * 10;  // d, vx, vy, vz, p, bx, by, bz, eint, gamma
            d = v[0];
            vx = v[1];
            vy = v[2];
            p = v[4];
            b1 = v[5];
            b2 = v[6];
            b3 = v[7];
        } else {
            double u[NCONS];
#pragma unroll
            for (int k = 0; k < NCONS; k++) u[k] = a.U[c * NCONS + k];
            const Phys ph{a.gamma, a.small_p, a.small_rho, a.hydro};
            const Prim P = cons2prim(ph, u);
            d = P.d;
            vx = P.vx;
            vy = P.vy;
            p = P.p;
            b1 = P.bx;
            b2 = P.by;
            b3 = P.bz;
        }
        unsigned long long x, y;
        cfl_speed_bits(a.gamma, a.hydro, d, vx, vy, p, b1, b2, b3, x, y);
        bx = x > bx ? x : bx;
        by = y > by ? y : by;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long wx = __shfl_xor(bx, o), wy = __shfl_xor(by, o);
        bx = wx > bx ? wx : bx;
        by = wy > by ? wy : by;
    }
    __shared__ unsigned long long sx[4], sy[4];
    if ((threadIdx.x & 63) == 0) {
        sx[threadIdx.x >> 6] = bx;
        sy[threadIdx.x >> 6] = by;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            bx = sx[w] > bx ? sx[w] : bx;
            by = sy[w] > by ? sy[w] : by;
        }
        part[2 * blockIdx.x] = bx;
        part[2 * blockIdx.x + 1] = by;
    }
}

// one block: the maxima of the first nblk partials, then dt as the host
// side of hmhd_cfl_dt forms it
__global__ __launch_bounds__(256) void k_cfl_noalloc_dt(const unsigned long long* __restrict__ part, int nblk,
                                                        double dx, double dy, double cfl, double* __restrict__ dt) {
    unsigned long long bx = 0, by = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        bx = part[2 * b] > bx ? part[2 * b] : bx;
        by = part[2 * b + 1] > by ? part[2 * b + 1] : by;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long wx = __shfl_xor(bx, o), wy = __shfl_xor(by, o);
        bx = wx > bx ? wx : bx;
        by = wy > by ? wy : by;
    }
    __shared__ unsigned long long sx[4], sy[4];
    if ((threadIdx.x & 63) == 0) {
        sx[threadIdx.x >> 6] = bx;
        sy[threadIdx.x >> 6] = by;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            bx = sx[w] > bx ? sx[w] : bx;
            by = sy[w] > by ? sy[w] : by;
        }
        const double xm = __longlong_as_double((long long)bx), ym = __longlong_as_double((long long)by);
        // dt = cfl * min(dx/xmax, dy/ymax) (src/cfl.f90:50); a NaN speed gives a NaN dt
        const double a = dx / xm, b = dy / ym;
        double m = a < b ? a : b;
        if (__builtin_isnan(a) || __builtin_isnan(b)) m = __builtin_nan("");
        *dt = cfl * m;
    }
}

// the message hmhd_last_error(NULL) returns (hmhd_api.cpp)
void set_free_error(const char* msg);

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned long long* g_scratch[kMaxDevices];  // scratch's device address, per device
}  // namespace

}  // namespace hmhd

extern "C" int hmhd_cfl_dt_noalloc(const hmhd_config* cfg, const double* U, const double* V, double* dt,
                                   void* stream) {
    using namespace hmhd;
    auto fail = [](int code, const char* msg) {
        set_free_error(msg);
        return code;
    };
    if (!cfg || !U || !dt) return fail(HMHD_ERR_ARG, "hmhd_cfl_dt_noalloc: null argument");
    if (cfg->R < 1 || cfg->R > kMaxR || cfg->nx < 1 || cfg->ny < 1 || !(cfg->xend > cfg->xbeg) ||
        !(cfg->yend > cfg->ybeg) || !(cfg->cfl > 0.0))
        return fail(HMHD_ERR_ARG, "hmhd_cfl_dt_noalloc: invalid configuration");
    const int ngc = cfg->R + 2;
    const long long mx = (long long)cfg->nx + 2 * ngc, my = (long long)cfg->ny + 2 * ngc;
    if (mx * my > HMHD_NOALLOC_MAX_CELLS)
        return fail(HMHD_ERR_ARG, "hmhd_cfl_dt_noalloc: more padded cells than HMHD_NOALLOC_MAX_CELLS");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(HMHD_ERR_DEVICE, "hmhd_cfl_dt_noalloc: no current device");
    unsigned long long* part;
    {
        // the array's address on this device, looked up once per device
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_scratch[dev]) {
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
                return fail(HMHD_ERR_DEVICE, "hmhd_cfl_dt_noalloc: hipGetSymbolAddress(scratch) failed");
            g_scratch[dev] = static_cast<unsigned long long*>(p);
        }
        part = g_scratch[dev];
    }
    NoallocCfl a;
    a.U = U;
    a.V = V;
    a.nx = cfg->nx;
    a.ny = cfg->ny;
    a.ngc = ngc;
    a.mx = (int)mx;
    a.hydro = cfg->force_hydro ? 1 : 0;
    a.gamma = cfg->gamma;
    a.small_p = cfg->small_p;
    a.small_rho = cfg->small_rho;
    const long long ncell = (long long)cfg->nx * cfg->ny;
    const int nblk = (int)((ncell + 255) / 256 < kNoallocBlocks ? (ncell + 255) / 256 : kNoallocBlocks);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_cfl_noalloc, dim3(nblk), dim3(256), 0, st, a, part);
    const double dx = (cfg->xend - cfg->xbeg) / cfg->nx, dy = (cfg->yend - cfg->ybeg) / cfg->ny;
    hipLaunchKernelGGL(k_cfl_noalloc_dt, dim3(1), dim3(256), 0, st, part, nblk, dx, dy, cfg->cfl, dt);
    if (hipGetLastError() != hipSuccess) return fail(HMHD_ERR_DEVICE, "hmhd_cfl_dt_noalloc: kernel launch failed");
    return HMHD_OK;
}
