// Shared host/device definitions of libnrhip (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "niftyreg_hip.h"

namespace nrh {

struct Error: std::runtime_error {
    int code;
    Error(int c, const std::string& m): std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define NRH_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::nrh::Error(NRH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define NRH_REQUIRE(cond, code, msg)                                  \
    do {                                                              \
        if (!(cond)) throw ::nrh::Error((code), std::string(msg));    \
    } while (0)

template<class F>
int guard(F&& f) {
    try {
        f();
        return NRH_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return NRH_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return NRH_ERR_ARG;
    }
}

// Owning device allocation.
template<class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(NRH_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    bool empty() const { return p == nullptr; }
    size_t bytes() const { return n * sizeof(T); }
};

// 3x4 affine (first three rows of a row-major 4x4), passed by value to kernels.
struct Affine {
    float m[12];
};

inline Affine to_affine(const float *m16) {
    Affine a;
    for (int i = 0; i < 12; ++i) a.m[i] = m16[i];
    return a;
}

// Geometry as seen by device code.
struct DevGeom {
    int nx, ny, nz;
    float dx, dy, dz;
    Affine xyz, ijk;
};

inline DevGeom to_dev(const nrh_geom& g) {
    DevGeom d;
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz;
    d.dx = g.dx; d.dy = g.dy; d.dz = g.dz;
    d.xyz = to_affine(g.xyz);
    d.ijk = to_affine(g.ijk);
    return d;
}

inline size_t voxels(const nrh_geom& g) { return (size_t)g.nx * g.ny * g.nz; }

void check_geom(const nrh_geom *g, const char *what);

// Device address of nrh_lts_static's static scratch on the current device (aladin.hip).
uint32_t *lts_static_scratch();

}  // namespace nrh
