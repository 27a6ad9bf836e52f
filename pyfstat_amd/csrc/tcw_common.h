// tcw_common.h -- host helpers shared by the translation units of libtcw_hip.so
// (tcw_map.hip: maps and merge; tcw_refine.hip: exponential-window refinement;
// tcw_demod.hip: F-statistic atoms from SFTs).
// Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "tcw_hip.h"

namespace tcwc {

// Message of the last failing call on this thread (tcw_last_error).
inline thread_local std::string g_err;

inline int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define TCW_HIP_TRY(expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::tcwc::fail(TCW_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Grow-only device buffer, or (view = true) a window into memory the buffer
// does not own, which it never frees or grows.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool view = false;
    void set_view(void *q, size_t bytes) {
        release();
        p = q;
        cap = bytes;
        view = true;
    }
    int ensure(size_t bytes) {
        if (bytes <= cap) return TCW_OK;
        if (view) return fail(TCW_ERR_INVALID, "internal: a scratch view cannot grow");
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 256);
        if (hipMalloc(&p, want) != hipSuccess) {
            p = nullptr;
            return fail(TCW_ERR_NOMEM, "hipMalloc of " + std::to_string(want) + " bytes failed");
        }
        cap = want;
        return TCW_OK;
    }
    void release() {
        if (p && !view) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        view = false;
    }
};

// Device of a handle (defined with the handle in tcw_map.hip).
int handle_device(const tcw_handle *h);

inline int set_device(int device) {
    int cur = -1;
    TCW_HIP_TRY(hipGetDevice(&cur));
    if (cur != device) TCW_HIP_TRY(hipSetDevice(device));
    return TCW_OK;
}

}  // namespace tcwc
