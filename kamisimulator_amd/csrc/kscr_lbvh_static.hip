// kscr_lbvh_static.hip -- the LBVH self query without device allocation.
//
// kscr_lbvh_self_query_static is kscr_lbvh_compute_device followed by
// kscr_lbvh_query_device, for callers that cannot allocate device memory (code
// under stream capture, serving loops, harnesses that forbid hipMalloc).  The
// tree, the sort buffers and the counters live in one file-scope device array
// of fixed size.  Its address on each device is looked up once, with that
// device current, the first time the entry point runs there.  Every call on
// a device shares that device's array and leaves the last call's contents in
// it: calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "kscr.h"
#include "kscr_lbvh.h"

using namespace kscr;

namespace {

// 128 Ki boxes: the reference's own 100k-box recipe fits; about 17 MiB of
// tree per device (lbvh_carve lays it out).
constexpr int kStaticMaxBoxes = 1 << 17;
constexpr size_t kScratchBytes = (size_t)24 << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

void set_err(char* err, size_t len, const char* msg) {
    if (err && len) {
        std::strncpy(err, msg, len - 1);
        err[len - 1] = 0;
    }
}

// The device address of `scratch` on the current device (looked up once per
// device).
hipError_t scratch_on(int device, char** out) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = p;
    }
    *out = static_cast<char*>(g_scratch[device]);
    return hipSuccess;
}

}  // namespace

extern "C" int kscr_lbvh_self_query_static(const float* d_aabbs, int n, int* d_pairs, size_t cap, size_t* count,
                                           void* stream, char* err, size_t err_len) {
    if (!count || (cap > 0 && !d_pairs) || cap > (size_t)INT32_MAX || (n > 0 && !d_aabbs) || n < 0) {
        set_err(err, err_len, "kscr_lbvh_self_query_static: invalid argument");
        return KSCR_ERR_INVALID_ARGUMENT;
    }
    if (n > kStaticMaxBoxes) {
        set_err(err, err_len, "kscr_lbvh_self_query_static: more than 131072 boxes (the static tree's capacity)");
        return KSCR_ERR_INVALID_ARGUMENT;
    }
    *count = 0;
    if (n == 0) return KSCR_OK;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) {
        set_err(err, err_len, "kscr_lbvh_self_query_static: no current device");
        return KSCR_ERR_HIP;
    }
    char* base = nullptr;
    if (scratch_on(device, &base) != hipSuccess) {
        set_err(err, err_len, "kscr_lbvh_self_query_static: hipGetSymbolAddress failed");
        return KSCR_ERR_HIP;
    }
    LbvhDev d;
    if (lbvh_carve(d, n, base, kScratchBytes) > kScratchBytes) {
        set_err(err, err_len, "kscr_lbvh_self_query_static: tree does not fit the static buffer");
        return KSCR_ERR_INVALID_ARGUMENT;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // the query's error word is the third word of the tree's counter block
    d.src = d_aabbs;  // the boxes are read in place
    if (lbvh_build(d, n, s) != hipSuccess ||
        lbvh_query(d, reinterpret_cast<int2*>(d_pairs), (int)cap, -1, d.count + 2, s) != hipSuccess) {
        set_err(err, err_len, "kscr_lbvh_self_query_static: kernel launch failed");
        return KSCR_ERR_HIP;
    }
    int cnt[3] = {0, 0, 0};  // pairs found, redo count, error word
    if (hipMemcpyAsync(cnt, d.count, sizeof cnt, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_err(err, err_len, "kscr_lbvh_self_query_static: copy of the pair count failed");
        return KSCR_ERR_HIP;
    }
    *count = (size_t)cnt[0];
    if (cnt[2]) {
        set_err(err, err_len, "kscr_lbvh_self_query_static: a query exceeded the traversal stack");
        return KSCR_ERR_MAX_COLLISIONS;
    }
    return KSCR_OK;
}
