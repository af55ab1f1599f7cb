// Internal helpers shared by the libazp translation units: error reporting,
// HIP call checking, device buffers and the orthorhombic periodic box.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "azp.h"

namespace azp
{

// Device buffer of at least N int4 owned by the launchers (one per device,
// kept until the process ends): fixed-point positions they pack themselves.
int launcher_heap_scratch(uint32_t N, int4** out);

// Whether the DPD kernel can stage the FP64 positions of a tiled list's
// neighbor set of `slots` entries in one workgroup's LDS.
bool dpd_tile_fits(uint32_t ntypes, uint32_t slots, uint32_t tile_size);

struct StepKernelArgs;

// The list side of a step's rebuild check (k_nlist.hip): the positions of
// the last build, the flag word the step reports into (cleared here on
// stream s) and the threshold; after the launch, which positions that flag
// checked.  nlist_update_checked is azp_nlist_update that reads such a check
// of p->d_pos instead of checking again.
int nlist_step_check(azp_nlist* nl, hipStream_t s, StepKernelArgs* k);
void nlist_step_checked(azp_nlist* nl, const double* d_pos);
int nlist_update_checked(azp_nlist* nl, const azp_particles* p, int force_rebuild, int* rebuilt, void* stream);

// Records a message for azp_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define AZP_HIP(call)                                                                          \
    do                                                                                         \
        {                                                                                      \
        hipError_t azp_e_ = (call);                                                            \
        if (azp_e_ != hipSuccess)                                                              \
            return ::azp::set_error(AZP_ERR_HIP,                                               \
                                    "%s failed: %s (%s:%d)",                                   \
                                    #call,                                                     \
                                    hipGetErrorString(azp_e_),                                 \
                                    __FILE__,                                                  \
                                    __LINE__);                                                 \
        } while (0)

#define AZP_TRY(call)                                                                          \
    do                                                                                         \
        {                                                                                      \
        int azp_s_ = (call);                                                                   \
        if (azp_s_ != AZP_OK)                                                                  \
            return azp_s_;                                                                     \
        } while (0)

#define AZP_LAUNCH_CHECK()                                                                     \
    do                                                                                         \
        {                                                                                      \
        hipError_t azp_e_ = hipGetLastError();                                                 \
        if (azp_e_ != hipSuccess)                                                              \
            return ::azp::set_error(AZP_ERR_HIP,                                               \
                                    "kernel launch failed: %s (%s:%d)",                        \
                                    hipGetErrorString(azp_e_),                                 \
                                    __FILE__,                                                  \
                                    __LINE__);                                                 \
        } while (0)

// Growable device allocation owned by a library object.
template<typename T> struct DeviceArray
    {
    T* ptr = nullptr;
    size_t count = 0;

    DeviceArray() = default;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    ~DeviceArray()
        {
        release();
        }

    void release()
        {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
        }

    // Ensures capacity for n elements (contents are not preserved on growth).
    int reserve(size_t n)
        {
        if (n <= count && ptr)
            return AZP_OK;
        release();
        size_t want = n > 0 ? n : 1;
        hipError_t e = hipMalloc((void**)&ptr, want * sizeof(T));
        if (e != hipSuccess)
            {
            ptr = nullptr;
            return set_error(AZP_ERR_NOMEM,
                             "hipMalloc of %zu bytes failed: %s",
                             want * sizeof(T),
                             hipGetErrorString(e));
            }
        count = want;
        return AZP_OK;
        }
    };

// Orthorhombic periodic box as the kernels see it.
struct DevBox
    {
    double L[3];
    double Linv[3];
    };

inline DevBox make_box(const azp_box& b)
    {
    DevBox d;
    d.L[0] = b.Lx;
    d.L[1] = b.Ly;
    d.L[2] = b.Lz;
    for (int k = 0; k < 3; ++k)
        d.Linv[k] = 1.0 / d.L[k];
    return d;
    }

inline bool box_valid(const azp_box& b)
    {
    return b.Lx > 0 && b.Ly > 0 && b.Lz > 0;
    }

// Round up to a multiple of the 64-lane wavefront.
inline uint32_t wave_pitch(uint32_t n)
    {
    return (n + 63u) & ~63u;
    }

inline unsigned grid_for(size_t n, unsigned block)
    {
    return (unsigned)((n + block - 1) / block);
    }

// The step after a pair force, with the next rebuild check (k_integrate.hip).
int launch_step(const double4* force, uint32_t N, const DevBox& box, const StepKernelArgs& step, hipStream_t s);

} // namespace azp
